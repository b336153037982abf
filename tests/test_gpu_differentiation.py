"""The differentiation configuration's finite-difference check (GPU).

Central differences with h = 1e-3 of an L2 loss on 256 randomly chosen
scalar parameters -- position, scale, rotation, beta, opacity and albedo,
about 43 of each -- against the full-gradient adjoint, at 10,000 clustered
surfels, 256 x 256, 256 samples per pixel, depth 8, with a fixed RNG (the
method of python/finite_difference/fd_test.py:97-298, with every scalar
type instead of opacity and beta only).

What is differentiated.  The forward renders over a frozen photon map
(render_view: the map is built once, as the adjoint treats it), so a
forward and the next are the same function of the parameters.  The
reference's adjoint estimator is the derivative of that function only where
three of its choices do not bite (AdjointKernels.cpp): completed Null events
are weighted by qNull (:183 after :274), a surfel endpoint enters with its
reflectance but not its own alpha*eta, and projection events exist only at
bounce 0 (:310), so a surfel behind another gets no direct term.  The
parameters are therefore drawn from surfels that are alone on every camera
ray through them (nothing in front, nothing behind, checked with 4 rays per
pixel), where the estimator and the forward agree in expectation.  Its
gradients elsewhere are pinned to the reference's estimator event by event
on the CPU (tests/test_oracle_transport_pins.py) and to the oracle on the
GPU (test_gpu_configs.py, test_gpu_parity.py).

Noise.  Both sides are Monte Carlo estimates with independent samples: the
forward's 256 passes are 8 groups of 32 seeded apart (the finite difference
of the image-weighted loss is linear in the image, so the 8 group
differences average to the 256-sample one and their spread gives its
standard error), and the adjoint is 8 independent 256-sample replicates
(adjoint_seed 42 + 1000 k, adjoint_decorrelate).  A parameter passes when
|adjoint - FD| <= 3 sigma + 5 % |FD| (sigma the combined standard error; 5 %
is the reference's relative bar, batch_fd_check.py:228-232, for what the
spread does not cover: the irradiance lookup moving with the hit point,
which neither estimator differentiates).  At most 4 % of the 256 may fail:
with sigma estimated from 8 + 8 replicates the 3-sigma test is a t-test
with about 14 degrees of freedom, which about 1 % of unbiased estimates
fail, and the heavy tails of the adjoint's path weights fail a few more.  Over all 256, the adjoint must
correlate with FD above 0.98, and the noise-weighted least-squares slope of
adjoint on FD (weights 1 / sigma^2) must lie within 10 % of 1.  (The plain
relative error of the whole vector is no bar: its largest entries are also
its noisiest.)
"""
import numpy as np
import pytest

from differentiablepointrendering_amd import scene

pytestmark = pytest.mark.gpu

N, W, SPP, G, R, H, NPARAM = 10_000, 256, 256, 8, 8, 1e-3, 256
TYPES = ("position", "scale", "rotation", "beta", "opacity", "albedo")


def _isolated_surfels(r, cam, cur, n):
    """Surfels (base index) that are the only surfel on every one of 4 rays
    per pixel that reaches them, with their ray counts."""
    inv = cam.inv_view.astype(np.float64)
    fy = 0.5 * cam.height / np.tan(np.radians(cam.fovy_deg) / 2)
    ys, xs = np.mgrid[0:cam.height, 0:cam.width]
    bad = np.zeros(n, bool)
    seen = np.zeros(n, np.int64)
    for ox, oy in ((0.25, 0.25), (0.75, 0.25), (0.25, 0.75), (0.75, 0.75)):
        d = np.stack([(xs + ox - cam.width / 2) / fy, (cam.height / 2 - ys - oy) / fy,
                      -np.ones(xs.shape)], -1).reshape(-1, 3)
        d = d @ inv[:3, :3].T
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        o = np.broadcast_to(inv[:3, 3], d.shape).astype(np.float32)
        d = d.astype(np.float32)
        h1 = r.trace_rays(o, d)
        hit = h1["instance"] >= 0
        cloud = np.bincount(h1["instance"][hit]).argmax()
        first = hit & (h1["instance"] == cloud)
        p1 = cur[h1["primitive"][first]]
        seen += np.bincount(p1, minlength=n)
        # what lies behind the first surfel, or behind the light's quad
        again = hit
        o2 = (h1["position"][again] + 1e-4 * d[again]).astype(np.float32)
        h2 = r.trace_rays(o2, d[again])
        behind = h2["instance"] >= 0
        first_of_again = first[again]
        bad[p1[behind[first_of_again]]] = True
        cl2 = behind & (h2["instance"] == cloud)
        bad[cur[h2["primitive"][cl2]]] = True
    return np.nonzero((seen >= 8) & ~bad)[0], seen


def test_c3_finite_differences_on_256_parameters():
    from differentiablepointrendering_amd.pale import Renderer

    pts, tgt = scene.clustered_surfels(N, seed=0)
    cam = scene.default_camera(W, W)
    meshes = [scene.light_quad()]
    base = dict(photons=1_000_000, bounces=8, forward_passes=1, gather_passes=SPP, adjoint_bounces=8,
                adjoint_passes=SPP, adjoint_decorrelate=1, seed=0, gather_radius=0.02, full_gradients=1)
    print(f"case: C3 finite differences, {N} clustered surfels, {W}x{W}, {SPP} spp, h={H}", flush=True)
    t = Renderer.from_scene(meshes, [], tgt, base)
    t.build_photon_map()
    target = t.render_view(cam, SPP, 5000)["raw"][..., :3].astype(np.float64)
    t.close()
    r = Renderer.from_scene(meshes, [], pts, base)
    r.build_photon_map()
    P = r.get_point_parameters()
    cur = np.arange(N)  # current BVH slot -> slot of the first build

    def set_params(Q):
        # apply_point_optimization + rebuild_bvh, as the reference's FD driver
        # rebuilds after every change (fd_test.py:137-160)
        nonlocal cur
        r.apply_point_optimization({k: Q[k][cur] for k in Q if k != "shape"})
        r.rebuild_bvh()
        cur = cur[r.surfel_bvh()["permutation"]]

    def images():
        return [r.render_view(cam, SPP // G, 100 + g * (SPP // G))["raw"][..., :3].astype(np.float64)
                for g in range(G)]

    set_params(P)
    img = sum(images()) / G
    dl = ((img - target) / img.size).astype(np.float32)
    # adjoint replicates: independent 256-sample estimates, in the first
    # build's slot order (each replica builds the same BVH from the same input)
    reps = []
    for k in range(R):
        q = Renderer.from_scene(meshes, [], pts, dict(base, adjoint_seed=42 + 1000 * k))
        q.build_photon_map()
        reps.append(q.render_view_backward(cam, SPP, dl))
        q.close()
    iso, seen = _isolated_surfels(r, cam, cur, N)
    print(f"c3: {len(iso)} isolated surfels of {int((seen > 0).sum())} visible", flush=True)
    assert len(iso) >= 64
    rng = np.random.default_rng(7)
    rows = []
    for j in range(NPARAM):
        ty = TYPES[j % len(TYPES)]
        s = int(rng.choice(iso))
        Qp = {k: v.copy() for k, v in P.items()}
        Qm = {k: v.copy() for k, v in P.items()}
        comp, axis = None, None
        if ty in ("position", "albedo"):
            comp = int(rng.integers(0, 3))
        elif ty == "scale":
            comp = int(rng.integers(0, 2))
        if ty == "rotation":
            axis = rng.normal(size=3)
            axis /= np.linalg.norm(axis)

            def rot(v, ang):
                v = v.astype(np.float64)
                c, sn = np.cos(ang), np.sin(ang)
                return (v * c + np.cross(axis, v) * sn + axis * (axis @ v) * (1 - c)).astype(np.float32)

            for Q, sg in ((Qp, 1.0), (Qm, -1.0)):
                Q["tangent_u"][s] = rot(P["tangent_u"][s], sg * H)
                Q["tangent_v"][s] = rot(P["tangent_v"][s], sg * H)
        elif comp is None:
            Qp[ty][s] += H
            Qm[ty][s] -= H
        else:
            Qp[ty][s, comp] += H
            Qm[ty][s, comp] -= H
        set_params(Qp)
        ip = images()
        set_params(Qm)
        im = images()
        # central difference of the loss, linearised in the image
        fd = np.array([float(np.sum(dl * (a - b))) / (2 * H) for a, b in zip(ip, im)])
        an = []
        for g in reps:
            if ty == "rotation":
                an.append(float(g["tangent_u"][s] @ np.cross(axis, P["tangent_u"][s]) +
                                g["tangent_v"][s] @ np.cross(axis, P["tangent_v"][s])))
            elif comp is None:
                an.append(float(g[ty][s]))
            else:
                an.append(float(g[ty][s, comp]))
        an = np.array(an)
        sigma = float(np.sqrt(fd.var(ddof=1) / G + an.var(ddof=1) / R))
        rows.append((ty, s, float(fd.mean()), float(an.mean()), sigma))
        if j % 32 == 0:
            print(f"c3: {j}/{NPARAM} {ty} fd {fd.mean():+.3e} adjoint {an.mean():+.3e} sigma {sigma:.1e}",
                  flush=True)
    set_params(P)
    r.close()
    fd = np.array([x[2] for x in rows])
    an = np.array([x[3] for x in rows])
    sig = np.array([x[4] for x in rows])
    ok = np.abs(an - fd) <= 3 * sig + 0.05 * np.abs(fd)
    for ty in TYPES:
        m = np.array([x[0] == ty for x in rows])
        print(f"c3: {ty}: {int((~ok[m]).sum())} of {int(m.sum())} outside the bar", flush=True)
    bad = [rows[i] for i in np.nonzero(~ok)[0]]
    corr = float(np.corrcoef(an, fd)[0, 1])
    w = np.where(sig > 0, 1.0 / np.maximum(sig, 1e-300) ** 2, 0.0)
    slope = float(np.sum(w * an * fd) / np.sum(w * fd * fd))
    z = (an - fd) / np.maximum(sig, 1e-300)
    print(f"c3: fail fraction {1 - ok.mean():.3f}, correlation {corr:.4f}, weighted slope {slope:.3f}, "
          f"median |z| {np.median(np.abs(z[sig > 0])):.2f}", flush=True)
    for x in rows:
        print("c3row", *x, flush=False)
    assert np.count_nonzero(fd) > NPARAM // 2
    assert 1 - ok.mean() <= 0.04, bad[:12]
    assert corr > 0.98, corr
    assert abs(slope - 1.0) <= 0.10, slope
