"""Serving of resident beta-surfel scenes on one GPU.

A ``SceneServer`` keeps scenes resident in device memory -- surfels, BVH and
photon map, built once per replica by ``Renderer.build_photon_map`` -- and
renders views of them on request, forward only or forward + adjoint against
a target image.  The reference has no serving layer; the per-view work is
its camera gather (PrimalKernels.cpp:500-653) and adjoint pass
(SyclBridge.cpp:169-294) over a resident photon map.

Scheduling is continuous: a request is started as soon as its scene has an
idle replica and its device memory fits the HBM budget, whatever else is
running; it never waits for a batch to fill.  Every replica is a renderer
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
            req, fut, reserve = job
            try:
                res = self._serve(req)
                ok = True
            except BaseException as e:  # reported through the future
                res, ok = e, False
            self.server._finished(self, reserve)
            if ok:
                fut.set_result(res)
            else:
                fut.set_exception(res)

    def _serve(self, req: ViewRequest) -> ViewResult:
        r = self.renderer
        started = time.perf_counter()
        img = r.render_view(req.camera, req.spp, req.seed)
        st = r.stats()
        queries = int(st["gather_queries"])
        device_ms = float(st["forward_ms"])
        grads = loss = None
        if req.target is not None:
            loss, dl = l2_adjoint(img["raw"][..., :3], req.target)
            grads = r.render_view_backward(req.camera, req.spp, dl)
            st = r.stats()
            queries += int(st["adjoint_queries"])
            device_ms += float(st["backward_ms"])
        return ViewResult(req, img["ldr"], img["raw"], grads, loss, queries, device_ms, started,
                          time.perf_counter(), self.index)


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
                 device_bytes: Optional[Callable[[], int]] = None):
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

    def _finished(self, rep: _Replica, reserve: int) -> None:
        with self._lock:
            rep.busy = False
            rep.last_used = time.perf_counter()
            rep.has_view_buffers = True
            self._reserved -= reserve
            self._running -= 1
            self._schedule()

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
            self._pending.remove(item)
            rep.busy = True
            self._reserved += need
            self._running += 1
            rep.inbox.append((req, fut, need))
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
