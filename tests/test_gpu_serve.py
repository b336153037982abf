"""Served views on the GPU: parity of a view and its adjoint with the oracle,
exact memory accounting of admission, and the same results alone and in a
mix of concurrent requests.

Bars: images within 1e-4 relative per pixel (observed bit-exact); opacity
gradients within 1e-4 relative plus an absolute floor of 1e-4 of the largest
gradient (FP32 atomics add in a different order), also between a request
served alone and in the mix.
"""
import numpy as np
import pytest

from differentiablepointrendering_amd import scene

pytestmark = pytest.mark.gpu

SMALL = dict(photons=20_000)


def _cam(eye, w, h, name="view"):
    return scene.look_at_camera(name, eye, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), w, h, 45.0)


@pytest.fixture(scope="module")
def pair():
    from differentiablepointrendering_amd.pale import Renderer
    from oracle.oracle import Oracle

    spec = scene.config_scene(0, n=4096, width=64, height=64)
    print("case: served views, config-0 scene n=4096", flush=True)
    gpu, orc = Renderer.from_spec(spec, SMALL), Oracle.from_spec(spec, SMALL)
    gpu.build_photon_map()
    orc.build_photon_map()
    return spec, gpu, orc


def test_view_without_photon_map_fails():
    from differentiablepointrendering_amd.pale import Renderer

    spec = scene.config_scene(0, n=256, width=16, height=16)
    r = Renderer.from_spec(spec, SMALL)
    with pytest.raises(RuntimeError, match="photon map|spp"):
        r.render_view(_cam((0, 0, 4), 8, 8), 1, 0)
    r.build_photon_map()
    with pytest.raises(RuntimeError, match="photon map|spp"):
        r.render_view(_cam((0, 0, 4), 8, 8), 0, 0)
    r.close()


def test_view_matches_oracle(pair):
    _, gpu, orc = pair
    cam = _cam((1.5, 1.0, 3.5), 48, 40)
    a = gpu.render_view(cam, 3, 777)
    b = orc.render_view(cam, 3, 777)
    np.testing.assert_allclose(a["raw"], b["raw"], rtol=1e-4, atol=1e-7)
    np.testing.assert_allclose(a["ldr"], b["ldr"], rtol=1e-4, atol=1e-6)
    assert a["raw"][..., :3].max() > 0
    # Another seed gives another image; the same seed the same one.
    c = gpu.render_view(cam, 3, 778)
    assert not np.array_equal(a["raw"], c["raw"])
    d = gpu.render_view(cam, 3, 777)
    assert np.array_equal(a["raw"].view(np.uint32), d["raw"].view(np.uint32))


def test_view_photon_map_matches_render_forward(pair):
    spec, gpu, orc = pair
    # build_photon_map and render_forward draw the same photon seeds.
    p_map = gpu.photons()
    p_orc = orc.photons()
    assert np.array_equal(p_map["position"].view(np.uint32), p_orc["position"].view(np.uint32))
    gpu.render_forward()
    p_fwd = gpu.photons()
    assert np.array_equal(p_map["position"].view(np.uint32), p_fwd["position"].view(np.uint32))
    gpu.build_photon_map()


def test_view_adjoint_matches_oracle(pair):
    _, gpu, orc = pair
    cam = _cam((-1.0, 0.5, 3.8), 40, 32)
    rng = np.random.default_rng(5)
    dl = (rng.normal(size=(32, 40, 3)) * 1e-3).astype(np.float32)
    g = gpu.render_view_backward(cam, 2, dl)
    go = orc.render_view_backward(cam, 2, dl)
    scale = np.abs(go).max()
    assert scale > 0
    np.testing.assert_allclose(g["opacity"], go, rtol=1e-4, atol=1e-4 * scale)


def test_view_memory_accounting_is_exact(pair):
    from differentiablepointrendering_amd.pale import Renderer

    _, gpu, _ = pair
    gpu.release_view_buffers()
    assert gpu.view_buffer_bytes() == 0
    cam = _cam((0.0, 2.0, 3.0), 96, 72)
    for spp, adjoint in ((1, False), (1, True), (5, False), (5, True)):
        want = gpu.view_bytes(96, 72, spp, adjoint)
        before = Renderer.device_bytes()
        gpu.render_view(cam, spp, 3)
        if adjoint:
            gpu.render_view_backward(cam, spp, np.zeros((72, 96, 3), np.float32))
        assert Renderer.device_bytes() - before == want, (spp, adjoint)
        assert gpu.view_bytes(96, 72, spp, adjoint) == 0
    held = gpu.view_buffer_bytes()
    img = gpu.render_view(cam, 2, 9)["raw"]
    before = Renderer.device_bytes()
    gpu.release_view_buffers()
    assert before - Renderer.device_bytes() == held
    assert np.array_equal(gpu.render_view(cam, 2, 9)["raw"], img)


def test_served_results_same_alone_and_in_mix():
    from differentiablepointrendering_amd.serve import SceneServer, ViewRequest

    specs = {"plain": scene.config_scene(0, n=4096), "clustered": scene.config_scene(1, n=20_000)}
    settings = {"plain": dict(specs["plain"].settings, photons=20_000),
                "clustered": dict(specs["clustered"].settings, photons=40_000, gather_radius=0.05)}
    rng = np.random.default_rng(1)
    reqs = []
    for i in range(12):
        name = "plain" if i % 2 == 0 else "clustered"
        w, h = int(rng.integers(24, 96)), int(rng.integers(24, 80))
        d = rng.normal(size=3)
        eye = tuple(4.0 * d / np.linalg.norm(d))
        target = (np.abs(rng.normal(size=(h, w, 3))) * 0.1).astype(np.float32) if i % 3 == 0 else None
        reqs.append(ViewRequest(name, _cam(eye, w, h), int(rng.choice([1, 2, 4])), 1000 + i, target=target))
    print(f"case: server alone vs mix, {len(reqs)} requests over 2 scenes x 2 replicas", flush=True)
    with SceneServer() as srv:
        for name, spec in specs.items():
            srv.add_scene(name, spec.meshes, spec.points, settings[name], replicas=2)
        alone = [srv.render(r) for r in reqs]
        mixed = [f.result(timeout=120) for f in [srv.submit(r) for r in reqs]]
    for a, m in zip(alone, mixed):
        assert np.array_equal(a.raw.view(np.uint32), m.raw.view(np.uint32))
        assert np.array_equal(a.ldr.view(np.uint32), m.ldr.view(np.uint32))
        if a.grads is not None:
            ga, gm = a.grads["opacity"], m.grads["opacity"]
            np.testing.assert_allclose(gm, ga, rtol=1e-4, atol=1e-4 * max(np.abs(ga).max(), 1e-30))
    assert sum(r.grads is not None for r in mixed) == 4


def test_scene_admission_rejects_beyond_budget():
    from differentiablepointrendering_amd.pale import Renderer
    from differentiablepointrendering_amd.serve import AdmissionError, SceneServer

    spec = scene.config_scene(0, n=4096)
    base = Renderer.device_bytes()
    with SceneServer(hbm_budget_bytes=base + (1 << 20)) as srv:
        with pytest.raises(AdmissionError):
            srv.add_scene("s", spec.meshes, spec.points, dict(spec.settings, photons=20_000))
        assert Renderer.device_bytes() == base


def test_views_memory_accounting_is_exact(pair):
    from differentiablepointrendering_amd.pale import Renderer

    _, gpu, _ = pair
    gpu.release_view_buffers()
    cams = [_cam((0.0, 2.0, 3.0), 40, 30, "a"), _cam((1.0, 1.0, 3.5), 64, 48, "b"), _cam((-1, 0, 3.8), 17, 9, "c")]
    spp = [3, 1, 2]
    adj = [True, False, True]
    want = gpu.views_bytes([(c.width, c.height) for c in cams], spp, adj)
    before = Renderer.device_bytes()
    out = gpu.render_views(cams, spp, [5, 6, 7])
    for c, s, a, o in zip(cams, spp, adj, out):
        if a:
            gpu.render_view_backward(c, s, np.zeros((c.height, c.width, 3), np.float32))
    assert Renderer.device_bytes() - before == want
    assert gpu.views_bytes([(c.width, c.height) for c in cams], spp, adj) == 0


def test_server_groups_small_requests():
    from differentiablepointrendering_amd.serve import SceneServer, ViewRequest

    spec = scene.config_scene(0, n=4096)
    reqs = [ViewRequest("s", _cam((0.5 * i, 1.0, 3.7), 20 + 3 * i, 16 + i), 1 + i % 3, 40 + i) for i in range(10)]
    print("case: grouped small requests", flush=True)
    with SceneServer(max_group=8) as srv:
        srv.add_scene("s", spec.meshes, spec.points, dict(spec.settings, photons=20_000))
        alone = [srv.render(r) for r in reqs]
        assert srv.groups_served == 0
        with srv._lock:  # hold the scheduler so that all ten wait together
            futs = [srv.submit(r) for r in reqs]
        mixed = [f.result(timeout=120) for f in futs]
        assert srv.groups_served >= 1
    for a, m in zip(alone, mixed):
        assert np.array_equal(a.raw.view(np.uint32), m.raw.view(np.uint32))
