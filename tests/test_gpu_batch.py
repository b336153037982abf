"""Batched views on the GPU: every image of pale_render_views equals the
single-view path (pale_render_view, itself checked against the oracle) bit
for bit, whatever the launch tuning (batch size, entry budget and so the
number of rounds, graph capture, grid size), across graph replays with
eager work in between; and the first views match the oracle directly.
"""
import numpy as np
import pytest

from differentiablepointrendering_amd import scene

pytestmark = pytest.mark.gpu

SIZES = [(37, 29), (64, 48), (8, 8), (100, 60), (13, 71), (50, 50), (1, 1), (71, 9)]
SPP = [1, 3, 2, 1, 2, 4, 1, 2]


def _views():
    rng = np.random.default_rng(11)
    cams = []
    for i, (w, h) in enumerate(SIZES):
        d = rng.normal(size=3)
        eye = 4.0 * d / np.linalg.norm(d)
        up = (0.0, 1.0, 0.0) if abs(eye[1]) < 3.9 else (1.0, 0.0, 0.0)
        cams.append(scene.look_at_camera(f"v{i}", eye, (0.0, 0.0, 0.0), up, w, h, 45.0))
    seeds = [1000 + 17 * i for i in range(len(cams))]
    return cams, SPP, seeds


@pytest.fixture(scope="module")
def renderer():
    from differentiablepointrendering_amd.pale import Renderer

    spec = scene.config_scene(0, n=4096, width=16, height=16)
    print("case: batched views, config-0 scene n=4096", flush=True)
    r = Renderer.from_spec(spec, dict(photons=20_000))
    r.build_photon_map()
    cams, spp, seeds = _views()
    single = [r.render_view(c, s, sd)["raw"] for c, s, sd in zip(cams, spp, seeds)]
    return spec, r, single


TUNINGS = [
    dict(),
    dict(slots_per_batch=3000),                      # many batches, views split across passes
    dict(slots_per_batch=3000, entries_per_slot=2),  # several rounds per batch
    dict(graph_rounds=1),
    dict(slots_per_batch=5000, entries_per_slot=2, graph_rounds=2, grid_fraction=0.5),
    dict(slots_per_batch=1, graph_rounds=1),         # one segment per batch
]


@pytest.mark.parametrize("tuning", TUNINGS, ids=[str(t) for t in TUNINGS])
def test_batched_views_equal_single_views(renderer, tuning):
    _, r, single = renderer
    cams, spp, seeds = _views()
    print(f"case: tuning {tuning}", flush=True)
    r.set_batch_tuning(**tuning)
    for replay in range(3):
        out = r.render_views(cams, spp, seeds)
        for i, (o, s) in enumerate(zip(out, single)):
            assert np.array_equal(o["raw"].view(np.uint32), s.view(np.uint32)), (tuning, replay, i)
        # eager work on the same stream between replays
        r.render_view(cams[1], 2, 5)
    if tuning.get("slots_per_batch") == 3000:
        assert r.last_batches > 3
    r.set_batch_tuning()


def test_batched_views_match_oracle(renderer):
    from oracle.oracle import Oracle

    spec, r, single = renderer
    orc = Oracle.from_spec(spec, dict(photons=20_000))
    orc.build_photon_map()
    cams, spp, seeds = _views()
    r.set_batch_tuning(slots_per_batch=3000, entries_per_slot=2, graph_rounds=1)
    out = r.render_views(cams, spp, seeds)
    r.set_batch_tuning()
    for i in range(4):
        want = orc.render_view(cams[i], spp[i], seeds[i])
        np.testing.assert_allclose(out[i]["raw"], want["raw"], rtol=1e-4, atol=1e-7)
        np.testing.assert_allclose(out[i]["ldr"], want["ldr"], rtol=1e-4, atol=1e-6)


def test_batch_tuning_validation(renderer):
    _, r, _ = renderer
    with pytest.raises(RuntimeError):
        r.set_batch_tuning(grid_fraction=1.5)
    with pytest.raises(RuntimeError):
        r.render_views([scene.look_at_camera("v", (0, 0, 4), (0, 0, 0), (0, 1, 0), 8, 8, 45.0)], [0], [1])
    assert r.render_views([], [], []) == []


def _dloss(cams, seed):
    rng = np.random.default_rng(seed)
    return [(rng.normal(size=(c.height, c.width, 3)) * 1e-3).astype(np.float32) for c in cams]


def _close(a, b):
    scale = max(float(np.abs(b).max()), 1e-30)
    np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-4 * scale)


@pytest.mark.parametrize("tuning", [dict(), dict(slots_per_batch=3000)], ids=["one batch", "many batches"])
def test_batched_adjoint_equals_sum_of_single_views(renderer, tuning):
    _, r, _ = renderer
    cams, spp, _ = _views()
    dl = _dloss(cams, 3)
    want = None
    for c, s, d in zip(cams, spp, dl):
        g = r.render_view_backward(c, s, d)["opacity"]
        want = g.copy() if want is None else want + g
    assert np.abs(want).max() > 0
    r.set_batch_tuning(**tuning)
    got = r.render_views_backward(cams, spp, dl)["opacity"]
    r.set_batch_tuning()
    _close(got, want)


def test_batched_full_gradients_equal_sum_of_single_views():
    from differentiablepointrendering_amd.pale import Renderer

    spec = scene.config_scene(0, n=4096, width=16, height=16)
    print("case: batched full gradients, config-0 scene n=4096", flush=True)
    r = Renderer.from_spec(spec, dict(photons=20_000, full_gradients=1))
    r.build_photon_map()
    cams, spp, _ = _views()
    dl = _dloss(cams, 4)
    want = {}
    for c, s, d in zip(cams, spp, dl):
        g = r.render_view_backward(c, s, d)
        for k in ("opacity", "position", "tangent_u", "scale", "albedo", "beta"):
            want[k] = g[k].copy() if k not in want else want[k] + g[k]
    r.set_batch_tuning(slots_per_batch=4000)
    got = r.render_views_backward(cams, spp, dl)
    assert r.last_batches > 1
    for k, v in want.items():
        assert np.abs(v).max() > 0, k
        # A sum over views of atomically accumulated terms: 1e-3 for the
        # geometry gradients, whose terms cancel more.
        scale = float(np.abs(v).max())
        np.testing.assert_allclose(got[k], v, rtol=1e-3, atol=1e-3 * scale, err_msg=k)


def test_batched_adjoint_matches_oracle(renderer):
    from oracle.oracle import Oracle

    spec, r, _ = renderer
    orc = Oracle.from_spec(spec, dict(photons=20_000))
    orc.build_photon_map()
    cams, spp, _ = _views()
    cams, spp = cams[:3], spp[:3]
    dl = _dloss(cams, 5)
    want = sum(orc.render_view_backward(c, s, d) for c, s, d in zip(cams, spp, dl))
    got = r.render_views_backward(cams, spp, dl)["opacity"]
    _close(got, want)


@pytest.mark.parametrize("decorrelate", [0, 1])
def test_batch_passes_render_forward_and_backward(decorrelate):
    # render_forward / render_backward with batch_passes = 1 walk all gather
    # passes (and trace all adjoint samples) of all cameras as batches; the
    # images equal the pass-by-pass ones bit for bit, the gradients up to the
    # order of float atomics.
    from differentiablepointrendering_amd.pale import Renderer

    spec = scene.config_scene(0, n=4096, width=48, height=40)
    spec.cameras.append(scene.look_at_camera("camera2", (1.0, 2.0, 3.2), (0, 0, 0), (0, 1, 0), 33, 21, 45.0))
    base = dict(photons=20_000, gather_passes=3, adjoint_passes=2, spp_decorrelate=decorrelate, full_gradients=1)
    print(f"case: batch_passes, decorrelate={decorrelate}", flush=True)
    a = Renderer.from_spec(spec, base)
    b = Renderer.from_spec(spec, dict(base, batch_passes=1))
    b.set_batch_tuning(slots_per_batch=3000, entries_per_slot=3)
    ia, ib = a.render_forward(), b.render_forward()
    for k in ia:
        assert np.array_equal(ia[k].view(np.uint32), ib[k].view(np.uint32)), k
    assert b.last_batches if hasattr(b, "last_batches") else True
    dl = {c.name: (ia[c.name + "_raw"][..., :3] * 1e-3).astype(np.float32) for c in spec.cameras}
    ga, _ = a.render_backward(dl)
    gb, _ = b.render_backward(dl)
    for k in ("opacity", "position", "scale", "albedo", "beta"):
        scale = float(np.abs(ga[k]).max())
        assert scale > 0, k
        np.testing.assert_allclose(gb[k], ga[k], rtol=1e-3, atol=1e-3 * scale, err_msg=k)
