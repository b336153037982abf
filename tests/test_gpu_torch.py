"""Device-memory entry points and the torch.autograd bridge (GPU)."""
import numpy as np
import pytest

from differentiablepointrendering_amd import scene

pytestmark = pytest.mark.gpu


def _renderer(spec, settings):
    from differentiablepointrendering_amd.pale import Renderer

    print(f"case: {spec.name} n={len(spec.points['position'])} settings={settings}", flush=True)
    return Renderer.from_spec(spec, settings)


def test_device_entry_points_match_host_ones():
    import torch

    spec = scene.config_scene(0, n=4096, width=64, height=64)
    r = _renderer(spec, dict(photons=20_000))
    host = r.render_forward()["camera1_raw"]
    dev = r.render_forward_device()["camera1_raw"]
    assert dev.is_cuda and dev.shape == (64, 64, 4)
    assert np.array_equal(dev.cpu().numpy().view(np.uint32), host.view(np.uint32))
    dl = (host[..., :3] / host[..., :3].size).astype(np.float32)
    gh, _ = r.render_backward({"camera1": dl})
    gd = r.render_backward_device({"camera1": torch.from_numpy(dl).cuda()})
    scale = np.abs(gh["opacity"]).max()
    assert scale > 0
    np.testing.assert_allclose(gd["opacity"].cpu().numpy(), gh["opacity"], rtol=1e-4, atol=1e-4 * scale)
    for k in ("position", "scale", "albedo", "beta"):
        assert not torch.any(gd[k]), k


def test_set_opacity_device_equals_host_update():
    import torch

    spec = scene.config_scene(0, n=4096, width=64, height=64)
    a = _renderer(spec, dict(photons=20_000))
    b = _renderer(spec, dict(photons=20_000))
    p = b.get_point_parameters()
    op = np.random.default_rng(4).uniform(0.1, 0.9, size=p["opacity"].shape).astype(np.float32)
    p["opacity"] = op
    b.apply_point_optimization(p)
    a.set_opacity_device(torch.from_numpy(op).cuda())
    assert np.array_equal(a.get_point_parameters()["opacity"], op)
    ia = a.render_forward()["camera1_raw"]
    ib = b.render_forward()["camera1_raw"]
    assert np.array_equal(ia.view(np.uint32), ib.view(np.uint32))


def test_autograd_opacity_gradient():
    import torch

    from differentiablepointrendering_amd.torch_ops import render_raw

    spec = scene.config_scene(0, n=4096, width=64, height=64)
    r = _renderer(spec, dict(photons=20_000))
    op0 = torch.from_numpy(r.get_point_parameters()["opacity"]).cuda()
    with torch.no_grad():
        target = 0.5 * render_raw(r, op0, "camera1")[..., :3]
    op = op0.clone().requires_grad_(True)
    img = render_raw(r, op, "camera1")
    loss = 0.5 * ((img[..., :3] - target) ** 2).mean()
    loss.backward()
    # the same step through the numpy entry points
    raw = r.render_forward()["camera1_raw"][..., :3]
    dl = ((raw - target.cpu().numpy()) / raw.size).astype(np.float32)
    g, _ = r.render_backward({"camera1": dl})
    scale = np.abs(g["opacity"]).max()
    assert scale > 0
    np.testing.assert_allclose(op.grad.cpu().numpy(), g["opacity"], rtol=1e-4, atol=1e-4 * scale)


def test_device_full_gradients_match_host():
    import torch

    spec = scene.config_scene(0, n=4096, width=64, height=64)
    r = _renderer(spec, dict(photons=20_000, full_gradients=1))
    raw = r.render_forward()["camera1_raw"][..., :3]
    dl = (raw / raw.size).astype(np.float32)
    gh, _ = r.render_backward({"camera1": dl})
    gd = r.render_backward_device({"camera1": torch.from_numpy(dl).cuda()})
    for k in ("position", "tangent_u", "scale", "albedo", "beta", "opacity"):
        m = np.abs(gh[k]).max()
        assert m > 0, k
        np.testing.assert_allclose(gd[k].cpu().numpy(), gh[k], rtol=1e-3, atol=1e-3 * m, err_msg=k)


def _tensors(r, requires_grad=True):
    import torch

    from differentiablepointrendering_amd.torch_ops import PARAMS

    p = r.get_point_parameters()
    return {k: torch.from_numpy(p[k]).cuda().requires_grad_(requires_grad) for k in PARAMS}


def test_render_surfels_gradients_are_the_adjoint():
    import torch

    from differentiablepointrendering_amd.torch_ops import render_surfels

    spec = scene.config_scene(0, n=4096, width=64, height=64)
    r = _renderer(spec, dict(photons=20_000, full_gradients=1))
    params = _tensors(r)
    with torch.no_grad():
        target = 0.8 * render_surfels(r, params, camera="camera1")[..., :3]
    img = render_surfels(r, params, camera="camera1")
    loss = 0.5 * ((img[..., :3] - target) ** 2).mean()
    loss.backward()
    raw = r.render_forward()["camera1_raw"][..., :3]
    dl = ((raw - target.cpu().numpy()) / raw.size).astype(np.float32)
    g, _ = r.render_backward({"camera1": dl})
    for k, t in params.items():
        want = g[k]
        scale = float(np.abs(want).max())
        assert scale > 0, k
        np.testing.assert_allclose(t.grad.cpu().numpy(), want, rtol=1e-3, atol=1e-3 * scale, err_msg=k)


def test_render_surfels_needs_full_gradients_for_geometry():
    from differentiablepointrendering_amd.torch_ops import render_surfels

    spec = scene.config_scene(0, n=256, width=16, height=16)
    r = _renderer(spec, dict(photons=5_000))
    params = _tensors(r, requires_grad=False)
    params["position"].requires_grad_(True)
    with pytest.raises(RuntimeError, match="full_gradients"):
        render_surfels(r, params, camera="camera1")


def test_render_surfels_frozen_map_finite_differences():
    # The reference's "empty" gradient scene (one anisotropic surfel with a
    # smooth rim), the photon map built once (render_view), central
    # differences of the loss against the autograd gradient for parameters
    # that keep the hit point in place, as the CPU oracle's frozen-map check.
    import os

    import torch

    from differentiablepointrendering_amd import assets
    from differentiablepointrendering_amd.pale import Renderer
    from differentiablepointrendering_amd.torch_ops import render_surfels

    root = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_scenes")
    meshes, cams = assets.load_xml(os.path.join(root, "GradientTests", "empty", "empty.xml"), root)
    ini = assets.load_points(os.path.join(root, "GradientTests", "empty", "initial.ply"))
    ini["opacity"][:] = 0.6
    ini["beta"][:] = 1.0
    ini["scale"][0] = [0.9, 0.6]
    tgt = {k: v.copy() for k, v in ini.items()}
    tgt["position"][0] += [0.05, 0.0, -0.04]
    tgt["scale"][0] *= 0.9
    tgt["beta"][0] += 0.5
    tgt["albedo"][0] = [0.5, 0.6, 0.7]
    base = dict(photons=100_000, bounces=5, forward_passes=5, seed=42, adjoint_bounces=1, full_gradients=1)
    cam = cams[0]
    view = (cam, 4, 7)
    print("case: frozen-map finite differences, reference scene 'empty'", flush=True)
    t = Renderer.from_scene(meshes, [], tgt, base)
    t.build_photon_map()
    target = torch.from_numpy(t.render_view(cam, 4, 7)["raw"][..., :3].copy()).cuda()
    t.close()
    r = Renderer.from_scene(meshes, [], ini, base)
    r.build_photon_map()
    p0 = _tensors(r, requires_grad=False)

    def loss(p):
        with torch.no_grad():
            img = render_surfels(r, p, view=view)
        return 0.5 * float(((img[..., :3] - target) ** 2).mean())

    params = {k: v.clone().requires_grad_(True) for k, v in p0.items()}
    img = render_surfels(r, params, view=view)
    (0.5 * ((img[..., :3] - target) ** 2).mean()).backward()
    g = {k: v.grad.cpu().numpy() for k, v in params.items()}
    tu, tv = p0["tangent_u"][0].clone(), p0["tangent_v"][0].clone()

    def moved(key, comp, dv):
        q = {k: v.clone() for k, v in p0.items()}
        if key == "rotation":
            c, s = np.cos(dv), np.sin(dv)
            q["tangent_u"][0] = c * tu + s * tv
            q["tangent_v"][0] = -s * tu + c * tv
        elif comp is None:
            q[key][0] += dv
        else:
            q[key][0, comp] += dv
        return q

    cases = [("opacity", None, 5e-3, 0.05), ("beta", None, 1e-2, 0.05), ("position", 0, 2e-3, 0.05),
             ("position", 2, 2e-3, 0.05), ("scale", 0, 2e-3, 0.05), ("scale", 1, 2e-3, 0.05),
             ("albedo", 1, 1e-2, 0.05), ("rotation", None, 2e-3, 0.3)]
    tu_n, tv_n = tu.cpu().numpy(), tv.cpu().numpy()
    for key, comp, h, tol in cases:
        fd = (loss(moved(key, comp, h)) - loss(moved(key, comp, -h))) / (2 * h)
        if key == "rotation":
            an = float(g["tangent_u"][0] @ tv_n - g["tangent_v"][0] @ tu_n)
        else:
            an = float(g[key][0] if comp is None else g[key][0, comp])
        print(key, comp, an, fd, flush=True)
        assert np.sign(an) == np.sign(fd), (key, comp, an, fd)
        assert abs(an - fd) <= tol * abs(fd), (key, comp, an, fd)


def test_apply_points_device_equals_host_apply():
    """pale_apply_points_device writes the records the host apply writes: the
    same images bit for bit, the same parameters read back, and the same BVH
    after a rebuild (the host mirror is refreshed from the device, also after
    two device applies in a row)."""
    import torch

    spec = scene.config_scene(0, n=4096, width=64, height=64)
    a = _renderer(spec, dict(photons=20_000))
    b = _renderer(spec, dict(photons=20_000))
    p = b.get_point_parameters()
    rng = np.random.default_rng(11)
    q = {k: v.copy() for k, v in p.items()}
    q["position"] = (p["position"] + rng.normal(0, 0.02, p["position"].shape)).astype(np.float32)
    q["tangent_u"] = (p["tangent_u"] + rng.normal(0, 0.05, p["tangent_u"].shape)).astype(np.float32)
    q["tangent_v"] = (p["tangent_v"] + rng.normal(0, 0.05, p["tangent_v"].shape)).astype(np.float32)
    q["scale"] = (p["scale"] * rng.uniform(0.8, 1.2, p["scale"].shape)).astype(np.float32)
    q["albedo"] = rng.uniform(0.1, 0.9, p["albedo"].shape).astype(np.float32)
    q["opacity"] = rng.uniform(0.1, 0.9, p["opacity"].shape).astype(np.float32)
    q["beta"] = rng.uniform(-1, 1, p["beta"].shape).astype(np.float32)
    b.apply_point_optimization(q)
    # two device applies in a row (the host mirror stays stale in between):
    # the second one is what counts
    z = {k: (v * np.float32(0.5)).astype(np.float32) for k, v in q.items() if k != "shape"}
    a.apply_point_optimization_device({k: torch.from_numpy(v).cuda() for k, v in z.items()})
    a.apply_point_optimization_device({k: torch.from_numpy(v).cuda() for k, v in q.items() if k != "shape"})
    ia = a.render_forward()["camera1_raw"]
    ib = b.render_forward()["camera1_raw"]
    assert np.array_equal(ia.view(np.uint32), ib.view(np.uint32))
    pa = a.get_point_parameters()
    for k in pa:
        assert np.array_equal(pa[k], b.get_point_parameters()[k]), k
    a.rebuild_bvh()
    b.rebuild_bvh()
    ba, bb = a.surfel_bvh(), b.surfel_bvh()
    for k in ("count", "left_first", "permutation"):
        assert np.array_equal(ba[k], bb[k]), k
    ia = a.render_forward()["camera1_raw"]
    ib = b.render_forward()["camera1_raw"]
    assert np.array_equal(ia.view(np.uint32), ib.view(np.uint32))
    with pytest.raises(RuntimeError):
        a.apply_point_optimization_device({k: torch.from_numpy(v[:10]).cuda() for k, v in q.items()})


def test_render_surfels_cpu_parameters_get_cpu_gradients():
    """Parameters on the host: the image is on the GPU and each gradient comes
    back on its parameter's device (autograd's requirement)."""
    import torch

    from differentiablepointrendering_amd.torch_ops import PARAMS, render_surfels

    spec = scene.config_scene(0, n=1024, width=32, height=32)
    r = _renderer(spec, dict(photons=10_000, full_gradients=1))
    host = r.get_point_parameters()
    params = {k: torch.from_numpy(host[k].copy()).requires_grad_(True) for k in PARAMS}
    img = render_surfels(r, params, camera="camera1")
    assert img.is_cuda
    (img[..., :3] ** 2).mean().backward()
    for k in PARAMS:
        assert params[k].grad is not None and params[k].grad.device.type == "cpu", k
    gd = {k: torch.from_numpy(host[k].copy()).cuda().requires_grad_(True) for k in PARAMS}
    img2 = render_surfels(r, gd, camera="camera1")
    assert torch.equal(img2, img)
    (img2[..., :3] ** 2).mean().backward()
    for k in PARAMS:
        np.testing.assert_allclose(gd[k].grad.cpu().numpy(), params[k].grad.numpy(), rtol=1e-4,
                                   atol=1e-4 * float(params[k].grad.abs().max()) + 1e-30, err_msg=k)
