"""Scheduling and HBM admission of the scene server, without a device.

The server takes its renderer factory and its device-memory probe as
parameters; here they are a stand-in whose memory use follows the same rules
as the HIP renderer (buffers grow to the largest view served, are released on
request and are freed with the renderer) and whose views take a set time, so
admission, continuous scheduling and overlap can be checked on the CPU.
"""
import threading
import time

import numpy as np
import pytest

from differentiablepointrendering_amd import scene
from differentiablepointrendering_amd.serve import AdmissionError, SceneServer, ViewRequest, latency_summary


class Memory:
    def __init__(self):
        self.lock = threading.Lock()
        self.bytes = 0
        self.running = 0
        self.max_running = 0

    def add(self, b):
        with self.lock:
            self.bytes += b

    def __call__(self):
        with self.lock:
            return self.bytes


class StandIn:
    """Renderer stand-in: 1000 bytes per surfel resident, 10 (+20 with the
    adjoint) bytes per pixel of the largest view served."""

    def __init__(self, mem: Memory, n: int, delay: float):
        self.mem, self.n, self.delay = mem, n, delay
        self.view = 0
        self.adj = 0
        self.closed = False
        self.last = {}
        self.group_calls = []
        mem.add(1000 * n)

    def build_photon_map(self):
        self.mem.add(500 * self.n)
        self.map = True

    def _need(self, w, h, adjoint):
        v = max(self.view, 10 * w * h)
        a = max(self.adj, 20 * w * h) if adjoint else self.adj
        return v, a

    def view_bytes(self, w, h, spp, adjoint):
        v, a = self._need(w, h, adjoint)
        return (v - self.view) + (a - self.adj)

    def _grow(self, w, h, adjoint):
        v, a = self._need(w, h, adjoint)
        self.mem.add((v - self.view) + (a - self.adj))
        self.view, self.adj = v, a

    def render_view(self, cam, spp, seed):
        self._grow(cam.width, cam.height, False)
        with self.mem.lock:
            self.mem.running += 1
            self.mem.max_running = max(self.mem.max_running, self.mem.running)
        time.sleep(self.delay)
        with self.mem.lock:
            self.mem.running -= 1
        v = np.float32((seed * 7919 + cam.width * 31 + spp) % 1000) / 1000
        img = np.full((cam.height, cam.width, 4), v, np.float32)
        self.last = dict(gather_queries=cam.width * cam.height * spp, forward_ms=self.delay * 1e3)
        return {"ldr": img, "raw": img.copy()}

    def render_view_backward(self, cam, spp, dl):
        self._grow(cam.width, cam.height, True)
        time.sleep(self.delay)
        self.last.update(adjoint_queries=cam.width * cam.height * spp, backward_ms=self.delay * 1e3)
        return {"opacity": np.full(self.n, float(dl.sum()), np.float32)}

    def render_views(self, cams, spp, seeds):
        self.group_calls.append(len(cams))
        total = sum(c.width * c.height for c in cams)
        grow = max(self.view, 10 * total) - self.view
        self.mem.add(grow)
        self.view += grow
        out = []
        for c, s, sd in zip(cams, spp, seeds):
            out.append(self.render_view(c, s, sd))
        self.last = dict(gather_queries=sum(c.width * c.height * s for c, s in zip(cams, spp)),
                         forward_ms=self.delay * 1e3)
        return out

    def views_bytes(self, sizes, spp, adjoint):
        # A group's views share one batch: its buffers hold all their pixels.
        v = max(self.view, 10 * sum(w * h for w, h in sizes))
        a = max([self.adj] + [20 * w * h for (w, h), ad in zip(sizes, adjoint) if ad])
        return (v - self.view) + (a - self.adj)

    def stats(self):
        return dict(self.last)

    def view_buffer_bytes(self):
        return self.view + self.adj

    def release_view_buffers(self):
        self.mem.add(-(self.view + self.adj))
        self.view = self.adj = 0

    def close(self):
        if not self.closed:
            self.mem.add(-(1000 * self.n + (500 * self.n if getattr(self, "map", False) else 0)
                           + self.view + self.adj))
            self.closed = True


def make_server(budget=None, delay=0.01, **kw):
    mem = Memory()
    made = []

    def factory(meshes, points, settings):
        r = StandIn(mem, len(points["position"]), delay)
        made.append(r)
        return r

    return SceneServer(budget, renderer_factory=factory, device_bytes=mem, **kw), mem, made


def pts(n):
    return {"position": np.zeros((n, 3), np.float32)}


def cam(w, h, name="v"):
    return scene.look_at_camera(name, (0, 0, 4), (0, 0, 0), (0, 1, 0), w, h, 45.0)


def test_results_same_alone_and_in_mix():
    srv, mem, _ = make_server(delay=0.005)
    with srv:
        srv.add_scene("a", [], pts(10), {}, replicas=2)
        srv.add_scene("b", [], pts(20), {}, replicas=2)
        reqs = [ViewRequest("ab"[i % 2], cam(8 + i, 6 + (i % 3)), 1 + i % 4, 100 + i,
                            target=np.zeros((6 + (i % 3), 8 + i, 3), np.float32) if i % 3 == 0 else None)
                for i in range(24)]
        alone = [srv.render(r) for r in reqs]
        futs = [srv.submit(r) for r in reqs]
        mixed = [f.result(timeout=30) for f in futs]
    for a, m in zip(alone, mixed):
        assert a.request is m.request
        assert np.array_equal(a.raw, m.raw)
        assert (a.grads is None) == (m.grads is None)
        if a.grads is not None:
            assert np.array_equal(a.grads["opacity"], m.grads["opacity"])
            assert a.loss == m.loss
    # Four replicas: up to four views at once.
    assert mem.max_running == 4


def test_scene_admission_by_budget_frees_on_reject():
    srv, mem, made = make_server(budget=100_000)
    with srv:
        assert srv.add_scene("a", [], pts(40), {}) == 60_000
        with pytest.raises(AdmissionError):
            srv.add_scene("b", [], pts(30), {})
        assert mem() == 60_000 and "b" not in srv.scenes
        assert made[-1].closed
    assert mem() == 0


def test_request_beyond_budget_is_rejected():
    srv, mem, _ = make_server(budget=100_000)
    with srv:
        srv.add_scene("a", [], pts(10), {})  # 15,000 resident
        f = srv.submit(ViewRequest("a", cam(100, 100), 1, 0))  # 100,000 more
        with pytest.raises(AdmissionError):
            f.result(timeout=10)
        ok = srv.render(ViewRequest("a", cam(50, 50), 1, 0))  # 25,000 more
        assert ok.raw.shape == (50, 50, 4)


def test_memory_blocked_request_waits_then_idle_buffers_are_released():
    srv, mem, made = make_server(budget=100_000, delay=0.05)
    with srv:
        srv.add_scene("a", [], pts(10), {})   # 15,000
        srv.add_scene("b", [], pts(10), {})   # 15,000
        # a: 40,000 of view buffers; b's request needs 40,000 more: waits for a.
        fa = srv.submit(ViewRequest("a", cam(80, 50), 1, 0))
        time.sleep(0.01)
        fb = srv.submit(ViewRequest("b", cam(80, 50), 1, 0))
        ra, rb = fa.result(timeout=10), fb.result(timeout=10)
        assert rb.started >= ra.finished
        # a's idle buffers were released to make room for b.
        assert made[0].view == 0 and made[1].view == 40_000
        assert mem() <= 100_000


def test_request_blocked_on_memory_holds_back_younger_ones():
    srv, mem, made = make_server(budget=100_000, delay=0.1)
    with srv:
        srv.add_scene("a", [], pts(10), {}, replicas=2)  # 30,000
        srv.add_scene("b", [], pts(10), {})              # 15,000
        r0 = srv.submit(ViewRequest("a", cam(50, 50), 1, 0))    # +25,000 -> 70,000: runs
        time.sleep(0.02)
        head = srv.submit(ViewRequest("a", cam(50, 70), 1, 1))  # +35,000 on the idle replica: blocked
        late = srv.submit(ViewRequest("b", cam(10, 10), 1, 2))  # +1,000 would fit, but is younger
        for f in (r0, head, late):
            f.result(timeout=10)
    assert late.result().started >= r0.result().finished
    assert late.result().started >= head.result().started
    # The head went to the replica whose buffers needed the least growth.
    assert head.result().replica == r0.result().replica


def test_requests_waiting_for_a_replica_do_not_hold_back_others():
    srv, mem, made = make_server(delay=0.1)
    with srv:
        srv.add_scene("a", [], pts(10), {})
        srv.add_scene("b", [], pts(10), {})
        fa = [srv.submit(ViewRequest("a", cam(8, 8), 1, i)) for i in range(3)]
        fb = srv.submit(ViewRequest("b", cam(8, 8), 1, 9))
        rb = fb.result(timeout=10)
        ra = [f.result(timeout=10) for f in fa]
    assert rb.started < ra[1].started


def test_close_cancels_pending_and_frees_memory():
    srv, mem, _ = make_server(delay=0.05)
    srv.add_scene("a", [], pts(10), {})
    futs = [srv.submit(ViewRequest("a", cam(8, 8), 1, i)) for i in range(5)]
    srv.close()
    assert any(f.cancelled() for f in futs)
    assert mem() == 0


def test_errors_are_reported_through_the_future():
    srv, mem, made = make_server()
    with srv:
        srv.add_scene("a", [], pts(4), {})
        made[0].render_view = lambda *a: (_ for _ in ()).throw(RuntimeError("device lost"))
        with pytest.raises(RuntimeError, match="device lost"):
            srv.render(ViewRequest("a", cam(4, 4), 1, 0))
        with pytest.raises(KeyError):
            srv.submit(ViewRequest("nope", cam(4, 4), 1, 0))
        with pytest.raises(ValueError):
            srv.submit(ViewRequest("a", cam(4, 4), 0, 0))


def test_latency_summary():
    srv, mem, _ = make_server(delay=0.002)
    with srv:
        srv.add_scene("a", [], pts(4), {}, replicas=2)
        t0 = time.perf_counter()
        res = [f.result() for f in [srv.submit(ViewRequest("a", cam(4, 4), 1, i)) for i in range(20)]]
        s = latency_summary(res, time.perf_counter() - t0)
    assert s["requests"] == 20 and s["p50_ms"] <= s["p99_ms"]
    assert s["queries"] == 20 * 16


def test_waiting_small_requests_of_a_scene_are_served_as_a_group():
    srv, mem, made = make_server(delay=0.05, max_group=4, group_slots=10_000)
    with srv:
        srv.add_scene("a", [], pts(10), {})
        first = srv.submit(ViewRequest("a", cam(20, 20), 1, 0))  # runs alone
        time.sleep(0.01)
        small = [srv.submit(ViewRequest("a", cam(10 + i, 10), 2, 10 + i,
                                        target=np.zeros((10, 10 + i, 3), np.float32) if i == 1 else None))
                 for i in range(5)]
        big = srv.submit(ViewRequest("a", cam(200, 100), 1, 99))  # too large to join a group
        res = [f.result(timeout=10) for f in [first] + small + [big]]
    assert made[0].group_calls == [4], made[0].group_calls   # 4 of the 5 small ones, then the fifth alone
    assert srv.groups_served == 1
    # Same results as alone.
    for r in res:
        want = np.float32((r.request.seed * 7919 + r.request.camera.width * 31 + r.request.spp) % 1000) / 1000
        assert np.all(r.raw == want)
    assert res[2].grads is not None and res[2].loss is not None
    assert sum(r.queries for r in res) == sum(r.request.camera.width * r.request.camera.height * r.request.spp * (2 if r.request.target is not None else 1) for r in res)


def test_group_is_admitted_by_its_memory():
    srv, mem, made = make_server(budget=100_000, delay=0.05, max_group=4, group_slots=20_000)
    with srv:
        srv.add_scene("a", [], pts(10), {})                        # 15,000
        first = srv.submit(ViewRequest("a", cam(50, 50), 1, 0))    # +25,000 of view buffers
        time.sleep(0.01)
        # One 80x80 view: 64,000 of buffers (79,000 in all) fits; the two
        # together would hold 128,000: the head goes alone, then the other.
        fs = [srv.submit(ViewRequest("a", cam(80, 80), 1, i)) for i in range(1, 3)]
        for f in [first] + fs:
            f.result(timeout=10)
    assert made[0].group_calls == [], made[0].group_calls
    assert mem() <= 100_000
