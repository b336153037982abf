"""Serving of resident beta-surfel scenes on one GPU.

A ``SceneServer`` keeps scenes resident in device memory -- surfels, BVH and
photon map, built once per replica by ``Renderer.build_photon_map`` -- and
renders views of them on request, forward only or forward + adjoint against
a target image.  The reference has no serving layer; the per-view work is
its camera gather (PrimalKernels.cpp:500-653) and adjoint pass
(SyclBridge.cpp:169-294) over a resident photon map.

Scheduling is continuous: a request is started as soon as its scene has an
idle replica and its device memory fits the HBM budget, whatever else is
running; it never waits for a batch to fill.  Small requests of one scene
that are waiting when a replica frees up go with it as a group: their
forward passes are one batched render_views call (its images equal the
single views' bit for bit), their adjoints follow one by one.  Every replica is a renderer
with its own HIP stream and its own host thread, so the work of views of
different scenes (and of several replicas of one scene) overlaps on the
device.  A view's image depends only on (scene, settings, camera, spp,
seed): no state is carried from one request to the next, so a result is the
same alone and in any mix (the adjoint's float atomics may add in another
order: gradients agree to rounding).

Memory admission: a scene is admitted if its replicas fit the budget after
their photon maps are built; a request is admitted when the growth of the
replica's image-sized buffers (``Renderer.view_bytes``) fits what the budget
has left, counting the growth promised to requests already running.  Image
buffers of idle replicas are a cache: when releasing some of them (least
recently used first) makes the request fit, they are released; otherwise the
request waits for running requests to finish, and one that cannot fit even
with nothing running is failed with ``AdmissionError``.  While the oldest
waiting request is blocked on memory no younger request starts, so a large
request is never starved by small ones.  (Requests that wait only for a busy
replica of their scene do not hold others back.)

``add_scene`` holds the scheduler while it builds a scene's photon maps:
requests already running continue, new ones start after it returns.
"""
from __future__ import annotations

import collections
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from ._abi import Camera


class AdmissionError(RuntimeError):
    """A scene or a request that does not fit the HBM budget."""


@dataclass
class ViewRequest:
    scene: str
    camera: Camera
    spp: int
    seed: int
    target: Optional[np.ndarray] = None  # (H,W,3): forward + adjoint of the L2 loss when given
    submitted: float = 0.0

    @property
    def adjoint(self) -> bool:
        return self.target is not None


@dataclass
class ViewResult:
    request: ViewRequest
    ldr: np.ndarray
    raw: np.ndarray
    grads: Optional[Dict[str, np.ndarray]]
    loss: Optional[float]
    queries: int          # closest-hit queries of the view (gather + adjoint)
    device_ms: float      # device time of the view's forward (+ adjoint)
    started: float
    finished: float
    replica: int

    @property
    def latency_ms(self) -> float:
        return (self.finished - self.request.submitted) * 1e3

    @property
    def wait_ms(self) -> float:
        return (self.started - self.request.submitted) * 1e3


def l2_adjoint(raw_rgb: np.ndarray, target: np.ndarray):
    """Loss 0.5 * mean((I - T)^2) per channel value and its dL/dI, the L2
    loss of python/losses.py:10-45."""
    d = (raw_rgb.astype(np.float32) - target.astype(np.float32))
    n = np.float32(d.size)
    return float(0.5 * np.sum(d.astype(np.float64) ** 2) / d.size), (d / n).astype(np.float32)


class _Replica:
    def __init__(self, server: "SceneServer", scene: str, index: int, renderer):
        self.server = server
        self.scene = scene
        self.index = index
        self.renderer = renderer
        self.busy = False
        self.last_used = 0.0
        self.has_view_buffers = False
        self.inbox: "collections.deque" = collections.deque()
        self.wake = threading.Condition(server._lock)
        self.thread = threading.Thread(target=self._run, name=f"replica-{scene}-{index}", daemon=True)
        self.thread.start()

    def _run(self):
        lock = self.server._lock
        while True:
            with lock:
                while not self.inbox:
                    self.wake.wait()
                job = self.inbox.popleft()
            if job is None:
                return
            group, reserve = job
            try:
                results = self._serve(group)
                errors = None
            except BaseException as e:  # reported through the futures
                results, errors = None, e
            self.server._finished(self, reserve, len(group))
            for k, (req, fut) in enumerate(group):
                if errors is None:
                    fut.set_result(results[k])
                else:
                    fut.set_exception(errors)

    def _serve(self, group) -> List[ViewResult]:
        """The forward passes of every request of the group in one batched
        render_views call (a group of one: render_view), then each adjoint.
        The group's forward queries and device time are booked on its first
        result."""
        r = self.renderer
        reqs = [g[0] for g in group]
        started = time.perf_counter()
        if len(reqs) == 1:
            imgs = [r.render_view(reqs[0].camera, reqs[0].spp, reqs[0].seed)]
        else:
            imgs = r.render_views([q.camera for q in reqs], [q.spp for q in reqs], [q.seed for q in reqs])
        st = r.stats()
        fwd_q, fwd_ms = int(st["gather_queries"]), float(st["forward_ms"])
        done = time.perf_counter()
        out = []
        for k, (req, img) in enumerate(zip(reqs, imgs)):
            queries, device_ms = (fwd_q, fwd_ms) if k == 0 else (0, 0.0)
            grads = loss = None
            finished = done
            if req.target is not None:
                loss, dl = l2_adjoint(img["raw"][..., :3], req.target)
                grads = r.render_view_backward(req.camera, req.spp, dl)
                st = r.stats()
                queries += int(st["adjoint_queries"])
                device_ms += float(st["backward_ms"])
                finished = time.perf_counter()
            out.append(ViewResult(req, img["ldr"], img["raw"], grads, loss, queries, device_ms, started, finished,
                                  self.index))
        return out


@dataclass
class _Scene:
    name: str
    replicas: List[_Replica] = field(default_factory=list)
    resident_bytes: int = 0


class SceneServer:
    """Resident scenes, HBM admission and continuous scheduling of views.

    ``renderer_factory(meshes, points, settings)`` makes one replica (a
    ``Renderer`` by default) and ``device_bytes()`` reads the device memory
    held by every renderer of the process; both are parameters so the
    scheduling can be exercised without a device.
    """

    def __init__(self, hbm_budget_bytes: Optional[int] = None,
                 renderer_factory: Optional[Callable] = None,
                 device_bytes: Optional[Callable[[], int]] = None,
                 max_group: int = 8, group_slots: int = 4 << 20, tuning: Optional[dict] = None):
        if renderer_factory is None or device_bytes is None:
            from .pale import Renderer

            renderer_factory = renderer_factory or (
                lambda meshes, points, settings: Renderer.from_scene(meshes, [], points, settings))
            device_bytes = device_bytes or Renderer.device_bytes
        self._factory = renderer_factory
        self._device_bytes = device_bytes
        self.budget = hbm_budget_bytes
        self._lock = threading.RLock()
        self._scenes: Dict[str, _Scene] = {}
        self._pending: "collections.deque" = collections.deque()
        self._reserved = 0
        self._running = 0
        self._closed = False
        # Small requests of one scene that wait together are served by one
        # batched call: at most max_group requests of at most group_slots
        # pixel samples (width x height x spp) each and in all.
        self.max_group = max(1, int(max_group))
        self.group_slots = int(group_slots)
        self.groups_served = 0
        # Launch shape of every replica (Renderer.set_batch_tuning keys): a
        # smaller slots_per_batch cuts a large view into more, shorter walk
        # launches, between which the device takes other requests' kernels;
        # grid_fraction leaves part of the GPU to them.  Results do not
        # depend on it.
        self.tuning = dict(tuning) if tuning else None

    # ------------------------------------------------------------- scenes --
    def add_scene(self, name: str, meshes, points: dict, settings: dict, replicas: int = 1) -> int:
        """Makes ``replicas`` renderers of the scene and builds their photon
        maps.  Returns the device bytes the scene holds; raises
        AdmissionError (and frees what it made) if the budget is exceeded."""
        if replicas < 1:
            raise ValueError("replicas must be >= 1")
        with self._lock:
            if name in self._scenes:
                raise ValueError(f"scene '{name}' is already resident")
            before = self._device_bytes()
            made = []
            try:
                for _ in range(replicas):
                    r = self._factory(meshes, points, settings)
                    made.append(r)
                    if self.tuning:
                        r.set_batch_tuning(**self.tuning)
                    r.build_photon_map()
                    if self.budget is not None and self._device_bytes() + self._reserved > self.budget:
                        raise AdmissionError(
                            f"scene '{name}': {self._device_bytes() - before} device bytes for {len(made)} "
                            f"replica(s) exceed the HBM budget ({self.budget} bytes, "
                            f"{before} already resident)")
            except BaseException:
                for r in made:
                    r.close()
                raise
            sc = _Scene(name, resident_bytes=self._device_bytes() - before)
            sc.replicas = [_Replica(self, name, i, r) for i, r in enumerate(made)]
            self._scenes[name] = sc
            return sc.resident_bytes

    def remove_scene(self, name: str) -> None:
        with self._lock:
            sc = self._scenes.get(name)
            if sc is None:
                raise KeyError(name)
            if any(r.busy for r in sc.replicas) or any(q[0].scene == name for q in self._pending):
                raise RuntimeError(f"scene '{name}' has requests in flight")
            del self._scenes[name]
            for rep in sc.replicas:
                rep.inbox.append(None)
                rep.wake.notify()
        for rep in sc.replicas:
            rep.thread.join()
            rep.renderer.close()

    @property
    def scenes(self) -> List[str]:
        return list(self._scenes)

    # ----------------------------------------------------------- requests --
    def submit(self, req: ViewRequest) -> Future:
        fut: Future = Future()
        with self._lock:
            if self._closed:
                raise RuntimeError("server is closed")
            if req.scene not in self._scenes:
                raise KeyError(f"no resident scene '{req.scene}'")
            if req.spp < 1:
                raise ValueError("spp must be >= 1")
            if req.target is not None and req.target.shape[:2] != (req.camera.height, req.camera.width):
                raise ValueError("target image does not match the camera resolution")
            req.submitted = time.perf_counter()
            self._pending.append((req, fut))
            self._schedule()
        return fut

    def render(self, req: ViewRequest) -> ViewResult:
        return self.submit(req).result()

    def _finished(self, rep: _Replica, reserve: int, count: int = 1) -> None:
        with self._lock:
            rep.busy = False
            rep.last_used = time.perf_counter()
            rep.has_view_buffers = True
            self._reserved -= reserve
            self._running -= count
            self._schedule()

    @staticmethod
    def _slots(req: ViewRequest) -> int:
        return req.camera.width * req.camera.height * req.spp

    def _companions(self, head: ViewRequest, item) -> list:
        """Younger waiting requests of head's scene to serve with it."""
        if self.max_group <= 1 or self._slots(head) > self.group_slots:
            return []
        out, total, after = [], self._slots(head), False
        for other in self._pending:
            if other is item:
                after = True
                continue
            if not after or len(out) + 1 >= self.max_group:
                continue
            req, fut = other
            if req.scene != head.scene or fut.cancelled():
                continue
            s = self._slots(req)
            if s > self.group_slots or total + s > self.group_slots:
                continue
            out.append(other)
            total += s
        return out

    def _make_room(self, keep: _Replica, need: int) -> bool:
        """True if ``need`` more bytes fit the budget, after releasing image
        buffers of idle replicas (least recently used first) when that
        suffices; nothing is released when it would not."""
        short = self._device_bytes() + self._reserved + need - self.budget
        if short <= 0:
            return True
        idle = [(r, r.renderer.view_buffer_bytes()) for sc in self._scenes.values() for r in sc.replicas
                if r is not keep and not r.busy and r.has_view_buffers]
        if sum(b for _, b in idle) < short:
            return False
        for r, b in sorted(idle, key=lambda x: x[0].last_used):
            r.renderer.release_view_buffers()
            r.has_view_buffers = False
            short -= b
            if short <= 0:
                break
        return self._device_bytes() + self._reserved + need <= self.budget

    def _schedule(self) -> None:
        """Starts every waiting request that can start now (lock held)."""
        for item in list(self._pending):
            req, fut = item
            if fut.cancelled():
                self._pending.remove(item)
                continue
            idle = [r for r in self._scenes[req.scene].replicas if not r.busy]
            if not idle:
                continue
            # The idle replica whose buffers need the least growth.
            need, rep = min(((r.renderer.view_bytes(req.camera.width, req.camera.height, req.spp, req.adjoint), r)
                             for r in idle), key=lambda x: (x[0], x[1].index))
            if self.budget is not None and not self._make_room(rep, need):
                if self._running == 0:
                    # Nothing running will free memory: it can never fit.
                    self._pending.remove(item)
                    fut.set_exception(AdmissionError(
                        f"view {req.camera.width}x{req.camera.height} of '{req.scene}' needs {need} more "
                        f"device bytes than the HBM budget ({self.budget}) leaves"))
                    continue
                return  # the oldest request blocked on memory holds back younger ones
            group = [item]
            companions = self._companions(req, item)
            if companions:
                cand = group + companions
                gneed = rep.renderer.views_bytes([(q.camera.width, q.camera.height) for q, _ in cand],
                                                 [q.spp for q, _ in cand], [q.adjoint for q, _ in cand])
                if self.budget is None or self._make_room(rep, gneed):
                    group, need = cand, gneed
            for g in group:
                self._pending.remove(g)
            if len(group) > 1:
                self.groups_served += 1
            rep.busy = True
            self._reserved += need
            self._running += len(group)
            rep.inbox.append((group, need))
            rep.wake.notify()

    # -------------------------------------------------------------- close --
    def close(self) -> None:
        with self._lock:
            self._closed = True
            for req, fut in self._pending:
                fut.cancel()
            self._pending.clear()
            reps = [r for sc in self._scenes.values() for r in sc.replicas]
            for rep in reps:
                rep.inbox.append(None)
                rep.wake.notify()
        for rep in reps:
            rep.thread.join()
            rep.renderer.close()
        self._scenes.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def latency_summary(results: List[ViewResult], makespan_s: float) -> dict:
    """p50/p99 latency (submission to completion), mean wait and throughput."""
    lat = np.array([r.latency_ms for r in results], np.float64)
    wait = np.array([r.wait_ms for r in results], np.float64)
    q = sum(r.queries for r in results)
    return {
        "requests": len(results),
        "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
        "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
        "mean_wait_ms": float(wait.mean()) if len(wait) else None,
        "requests_per_s": len(results) / makespan_s if makespan_s > 0 else None,
        "mrays_per_s": q / makespan_s / 1e6 if makespan_s > 0 else None,
        "queries": int(q),
    }
