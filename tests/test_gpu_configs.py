"""The benchmark configurations at their full sizes on the GPU.

  C0  4,096 uniform surfels, 128x128, 4 spp, depth 4, 100k photons, r 0.05:
      everything against the oracle (BVH, photon map, image bit for bit;
      opacity and full gradients within tolerance).
  C1  1,048,576 clustered surfels, 800x800, 16 spp, depth 6, 2M photons,
      r 0.02: BVH nodes and permutation bit for bit, the photon map, a
      strided subset of the image pixels bit for bit (the oracle's bounded
      sample: pixels are independent given the photon map), opacity and full
      gradients of the whole 16-sample adjoint within tolerance.
  C2  32 views of 500k clustered surfels (eyes on the radius-4 sphere,
      128x96..1024x768, spp in {4, 8, 16}): one batched call = the views one
      by one, images bit for bit and gradient sums within tolerance; two
      views against the oracle on strided pixels, the smallest view's adjoint
      against the oracle.
  C4  the four resident scenes (100k, 500k, 1M, 3M clustered surfels with
      1-4M photon maps) serving the head of the seed-1 request stream: every
      result the same alone and in the mix.
(C3, the finite-difference check of the differentiation configuration, is
in test_gpu_differentiation.py.)

Bars: bit-exact for BVH, hits and images; photon positions bit for bit up
to a rare flip of a final rounding (> 99.9 % equal); gradients within 1e-4
of the largest entry for opacity and 1e-3 for the other parameters (FP32
atomics add in another order).
"""
import numpy as np
import pytest

from differentiablepointrendering_amd import scene

pytestmark = pytest.mark.gpu


def _pair(spec, settings=None):
    from differentiablepointrendering_amd.pale import Renderer
    from oracle.oracle import Oracle

    print(f"case: {spec.name} n={len(spec.points['position'])} settings={settings}", flush=True)
    return Renderer.from_spec(spec, settings), Oracle.from_spec(spec, settings)


def _photons_match(a, b):
    assert a["count"] == b["count"] > 0 and a["attempted"] == b["attempted"]
    same = np.all(a["position"] == b["position"], axis=1) & np.all(a["direction"] == b["direction"], axis=1)
    assert same.mean() > 0.999, same.mean()
    np.testing.assert_allclose(a["power"], b["power"], rtol=1e-5)


def _bvh_match(gpu, orc):
    a, b = gpu.surfel_bvh(), orc.surfel_bvh()
    for k in ("count", "left_first", "permutation"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["bounds"].view(np.uint32), b["bounds"].view(np.uint32))
    return a


def _grads_match(g, go, full):
    scale = np.abs(go).max()
    assert scale > 0
    np.testing.assert_allclose(g["opacity"], go, rtol=1e-4, atol=1e-4 * scale)
    for k in ("position", "tangent_u", "tangent_v", "scale", "albedo", "beta"):
        m = np.abs(full[k]).max()
        assert m > 0, k
        np.testing.assert_allclose(g[k], full[k], rtol=1e-3, atol=1e-3 * m, err_msg=k)


def _dloss(raw, seed):
    """dL/dI of an L2 loss against a perturbed copy of the image."""
    rng = np.random.default_rng(seed)
    t = raw[..., :3] * rng.uniform(0.5, 1.5, raw[..., :3].shape).astype(np.float32)
    return ((raw[..., :3] - t) / raw[..., :3].size).astype(np.float32)


# ---------------------------------------------------------------- C0 ----
def test_c0_full_config_against_oracle():
    spec = scene.config_scene(0)
    assert len(spec.points["position"]) == 4096 and spec.cameras[0].width == 128
    gpu, orc = _pair(spec, dict(full_gradients=1))
    _bvh_match(gpu, orc)
    ia, ib = gpu.render_forward(), orc.render_forward()
    _photons_match(gpu.photons(), orc.photons())
    assert gpu.stats()["photon_queries"] >= 100_000  # every emitted photon traced at least once
    for k in ib:
        assert np.array_equal(ia[k].view(np.uint32), ib[k].view(np.uint32)), k
    dl = _dloss(ib["camera1_raw"], 1)
    g, _ = gpu.render_backward({"camera1": dl})
    go = orc.render_backward({"camera1": dl})
    _grads_match(g, go, orc.full_gradients())
    a, b = gpu.stats(), orc.stats()
    for k in ("photon_queries", "gather_queries", "adjoint_queries", "gradient_events"):
        assert a[k] == b[k], k


# ---------------------------------------------------------------- C1 ----
C1_STRIDE = 61  # pixels with index % 61 == 0: about 10,500 of 640,000


@pytest.fixture(scope="module")
def c1():
    from oracle import oracle as orc_mod

    orc_mod.set_threads(0)
    spec = scene.config_scene(1)
    assert len(spec.points["position"]) == 1_048_576 and spec.cameras[0].width == 800
    gpu, orc = _pair(spec, dict(full_gradients=1))
    orc.set_sample(1, C1_STRIDE, 0)
    ia = gpu.render_forward()
    ib = orc.render_forward()
    print("c1: forward done", flush=True)
    return spec, gpu, orc, ia, ib


def test_c1_bvh_photon_map_and_image(c1):
    spec, gpu, orc, ia, ib = c1
    b = _bvh_match(gpu, orc)
    assert len(b["count"]) == 2 * (1_048_576 // 4) - 1
    _photons_match(gpu.photons(), orc.photons())
    a, o = ia["camera1_raw"].reshape(-1, 4), ib["camera1_raw"].reshape(-1, 4)
    idx = np.arange(0, a.shape[0], C1_STRIDE)
    assert np.array_equal(a[idx].view(np.uint32), o[idx].view(np.uint32))
    assert (o[idx, :3].sum(1) > 0).mean() > 0.2
    assert np.all(o[np.setdiff1d(np.arange(a.shape[0]), idx)] == 0)  # the oracle drew only the sample
    al, ol = ia["camera1"].reshape(-1, 4), ib["camera1"].reshape(-1, 4)
    assert np.array_equal(al[idx].view(np.uint32), ol[idx].view(np.uint32))


def test_c1_gradients_of_the_16_sample_adjoint(c1):
    spec, gpu, orc, ia, _ = c1
    dl = _dloss(ia["camera1_raw"], 2)
    g, _ = gpu.render_backward({"camera1": dl})
    orc.set_sample(1, 1, 0)  # the adjoint over every pixel
    go = orc.render_backward({"camera1": dl})
    print("c1: backward done", flush=True)
    _grads_match(g, go, orc.full_gradients())
    assert gpu.stats()["adjoint_queries"] == orc.stats()["adjoint_queries"]
    assert gpu.stats()["gradient_events"] == orc.stats()["gradient_events"]
    gpu.close()
    orc.close()


# ---------------------------------------------------------------- C2 ----
def test_c2_batched_views_equal_single_views_and_oracle():
    from differentiablepointrendering_amd.pale import Renderer
    from oracle.oracle import Oracle

    pts, _ = scene.clustered_surfels(500_000, seed=0)
    settings = dict(photons=2_000_000, bounces=6, forward_passes=1, gather_passes=16, adjoint_bounces=6,
                    adjoint_passes=16, seed=0, gather_radius=0.02, full_gradients=1)
    cams, spp = scene.batched_cameras(32)
    seeds = [100 + i for i in range(len(cams))]
    print(f"case: C2 32 views, {sum(c.width * c.height * s for c, s in zip(cams, spp)) / 1e6:.1f}M samples",
          flush=True)
    meshes = [scene.light_quad()]
    r = Renderer.from_scene(meshes, [], pts, settings)
    r.build_photon_map()
    single = [r.render_view(c, s, sd)["raw"] for c, s, sd in zip(cams, spp, seeds)]
    batched = r.render_views(cams, spp, seeds)
    for k, (a, b) in enumerate(zip(single, batched)):
        assert np.array_equal(a.view(np.uint32), b["raw"].view(np.uint32)), k
    dls = [_dloss(a, 10 + k) for k, a in enumerate(single)]
    keys = ("opacity", "position", "tangent_u", "tangent_v", "scale", "albedo", "beta")
    total = {k: 0.0 for k in keys}
    for c, s, d in zip(cams, spp, dls):
        g = r.render_view_backward(c, s, d)
        for k in keys:
            total[k] = total[k] + g[k]
    gb = r.render_views_backward(cams, spp, dls)
    for k in keys:
        m = max(float(np.abs(total[k]).max()), 1e-30)
        np.testing.assert_allclose(gb[k], total[k], rtol=1e-4 if k == "opacity" else 1e-3,
                                   atol=(1e-4 if k == "opacity" else 1e-3) * m, err_msg=k)
    # The oracle: the photon map, two views on strided pixels, the smallest
    # view's adjoint over all its pixels.
    orc = Oracle(meshes, [], pts, settings)
    orc.build_photon_map()
    _photons_match(r.photons(), orc.photons())
    order = np.argsort([c.width * c.height for c in cams])
    for v in (int(order[-1]), int(order[len(order) // 2])):
        orc.set_sample(1, 97, 0)
        o = orc.render_view(cams[v], spp[v], seeds[v])["raw"].reshape(-1, 4)
        a = single[v].reshape(-1, 4)
        idx = np.arange(0, a.shape[0], 97)
        assert np.array_equal(a[idx].view(np.uint32), o[idx].view(np.uint32)), v
    orc.set_sample(1, 1, 0)
    v = int(order[0])
    go = orc.render_view_backward(cams[v], spp[v], dls[v])
    g = r.render_view_backward(cams[v], spp[v], dls[v])
    _grads_match(g, go, orc.full_gradients())
    r.close()
    orc.close()


# ---------------------------------------------------------------- C4 ----
def test_c4_resident_scenes_same_alone_and_in_mix():
    from differentiablepointrendering_amd.serve import SceneServer, ViewRequest

    stream = scene.served_requests(12, 1.0)
    print(f"case: C4 4 resident scenes, {len(stream)} requests, "
          f"{sum(c.width * c.height * s for _, _, c, s, _ in stream) / 1e6:.0f}M samples", flush=True)
    rng = np.random.default_rng(3)
    with SceneServer() as srv:
        for k, (n, photons) in enumerate(scene.SERVED_SCENES):
            pts, _ = scene.clustered_surfels(n, seed=k)
            settings = dict(photons=photons, bounces=6, forward_passes=1, gather_passes=16, adjoint_bounces=6,
                            adjoint_passes=16, seed=0, gather_radius=0.02)
            srv.add_scene(f"s{k}", [scene.light_quad()], pts, settings, replicas=2)
        print("c4: scenes resident", flush=True)
        reqs = []
        for i, (_, sc, cam, spp, adj) in enumerate(stream):
            tgt = (np.abs(rng.normal(size=(cam.height, cam.width, 3))) * 0.05).astype(np.float32) if adj else None
            reqs.append(ViewRequest(f"s{sc}", cam, spp, 1000 + i, target=tgt))
        alone = [srv.render(q) for q in reqs]
        print("c4: alone done", flush=True)
        mixed = [f.result(timeout=300) for f in [srv.submit(q) for q in reqs]]
    assert any(q.target is not None for q in reqs) and len({q.scene for q in reqs}) == 4
    for a, m in zip(alone, mixed):
        assert np.array_equal(a.raw.view(np.uint32), m.raw.view(np.uint32))
        assert np.array_equal(a.ldr.view(np.uint32), m.ldr.view(np.uint32))
        if a.grads is not None:
            ga, gm = a.grads["opacity"], m.grads["opacity"]
            np.testing.assert_allclose(gm, ga, rtol=1e-4, atol=1e-4 * max(np.abs(ga).max(), 1e-30))
