"""Pins of the CPU oracle (CPU only).

The oracle is the checker of the HIP path, so it is itself checked here
against things that do not come from the oracle's own formulas:
published known-answer values (RNGs), closed forms (ray/plane and
ray/triangle distances, the beta kernel, compositing, post-processing),
brute force over all primitives / photons, and BVH invariants.  A
plausible restatement mistake (a dropped term, a wrong sign or index, a
transposed operand) fails at least one of them.
"""
import ctypes as C
import math

import numpy as np
import pytest

from differentiablepointrendering_amd import scene
from oracle import oracle as orc_mod
from oracle.oracle import Oracle

L = orc_mod.lib()
L.orc_splitmix64.argtypes = [C.c_uint64, C.c_uint32, C.POINTER(C.c_uint64)]
L.orc_xorshift128_state.argtypes = [C.c_uint32] * 5 + [C.POINTER(C.c_uint32)]
L.orc_item_rng.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.POINTER(C.c_float)]
L.orc_mt19937_64.argtypes = [C.c_uint64, C.c_uint32, C.POINTER(C.c_uint64)]


# ------------------------------------------------------------------- RNG --
def test_splitmix64_published_vector():
    # Vigna's reference splitmix64.c seeded with 1234567.
    out = (C.c_uint64 * 5)()
    L.orc_splitmix64(1234567, 5, out)
    assert list(out) == [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
                         16408922859458223821]


def test_xorshift128_published_vector():
    # Marsaglia, "Xorshift RNGs" (2003), xor128 with its stated initial state.
    out = (C.c_uint32 * 5)()
    L.orc_xorshift128_state(123456789, 362436069, 521288629, 88675123, 5, out)
    assert list(out) == [3701687786, 458299110, 2500872618, 3633119408, 516391518]


def test_mt19937_64_standard_value():
    # C++ [rand.predef]: the 10000th invocation of a default-constructed
    # mt19937_64 (seed 5489) produces 9981545732273789042.
    out = (C.c_uint64 * 10000)()
    L.orc_mt19937_64(5489, 10000, out)
    assert out[9999] == 9981545732273789042


def _splitmix(seed, n):
    M = (1 << 64) - 1
    out, s = [], seed
    for _ in range(n):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        out.append(z ^ (z >> 31))
    return out


def test_per_item_stream_composition():
    # makePerItemSeed1D + Xorshift128(uint64) + nextFloat
    # (KernelHelpers.h:79-115) composed from the pinned primitives.
    M = (1 << 64) - 1
    for base, item in [(42, 0), (42, 7), (0xDEADBEEF, 123456), (3, 2**31 + 5)]:
        seed = _splitmix((base ^ ((0x9E3779B97F4A7C15 + item) & M)) & M, 1)[0]
        st = [v & 0xFFFFFFFF for v in _splitmix(seed, 4)]
        x, y, z, w = st
        want = []
        for _ in range(6):
            t = (x ^ (x << 11)) & 0xFFFFFFFF
            x, y, z = y, z, w
            w = (w ^ (w >> 19)) ^ (t ^ (t >> 8))
            want.append(np.float32((w >> 8) * 2.0 ** -24))
        got = (C.c_float * 6)()
        L.orc_item_rng(base, item, 6, got)
        assert [np.float32(v) for v in got] == want


# -------------------------------------------------------- intersections --
def _one_surfel(beta=0.0, su=0.5, sv=0.5, opacity=0.7, pos=(0.0, 0.0, 0.0), tu=(1.0, 0.0, 0.0), tv=(0.0, 1.0, 0.0)):
    return scene.single_surfel_scene(position=pos, tangent_u=tu, tangent_v=tv, scale=(su, sv), opacity=opacity,
                                     beta=beta, width=16, height=16)


@pytest.mark.parametrize("beta", [0.0, -1.0, 0.7])
def test_surfel_hit_closed_form(beta):
    o = Oracle.from_spec(_one_surfel(beta=beta), dict(photons=1000))
    orig = np.array([[0.1, 0.2, 4.0], [0.0, 0.0, 2.0], [0.3, -0.1, -3.0]], np.float32)
    dirs = np.array([[0, 0, -1.0], [0, 0, -1.0], [0, 0, 1.0]], np.float32)
    h = o.trace(orig, dirs)
    assert list(h["instance"]) == [1, 1, 1]  # the cloud is the instance after the light mesh
    assert list(h["primitive"]) == [0, 0, 0]
    np.testing.assert_allclose(h["t"], [4.0, 2.0, 3.0], rtol=1e-6)
    np.testing.assert_allclose(h["position"][:, :2], orig[:, :2], atol=1e-6)
    r2 = (orig[:, 0] / 0.5) ** 2 + (orig[:, 1] / 0.5) ** 2
    want = (1.0 - r2) ** (4.0 * math.exp(beta))  # alpha = (1 - r^2)^(4 e^beta)
    np.testing.assert_allclose(h["alpha"], want, rtol=2e-6)


def test_surfel_support_edge_and_anisotropy():
    o = Oracle.from_spec(_one_surfel(su=0.2, sv=0.6), dict(photons=1000))
    # r^2 = (x/su)^2 + (y/sv)^2: inside along v, outside along u.
    orig = np.array([[0.0, 0.59, 1.0], [0.21, 0.0, 1.0], [0.19, 0.0, 1.0], [0.0, 0.61, 1.0]], np.float32)
    dirs = np.tile(np.array([[0, 0, -1.0]], np.float32), (4, 1))
    h = o.trace(orig, dirs)
    assert list(h["instance"] >= 0) == [True, False, True, False]
    assert h["alpha"][0] == pytest.approx((1 - (0.59 / 0.6) ** 2) ** 4, rel=1e-4)


def test_grazing_and_backward_rays_miss():
    o = Oracle.from_spec(_one_surfel(), dict(photons=1000))
    orig = np.array([[0.0, 0.0, 1.0], [-2.0, 0.0, 0.0], [0.0, 0.0, -1.0]], np.float32)
    dirs = np.array([[0, 0, 1.0], [1.0, 0, 0], [0, 0, -1.0]], np.float32)  # away, in-plane, away
    h = o.trace(orig, dirs)
    assert list(h["instance"]) == [-1, -1, -1]


def test_light_triangle_closed_form():
    o = Oracle.from_spec(_one_surfel(pos=(5.0, 5.0, 5.0)), dict(photons=1000))
    # The light quad spans x,z in [-0.25, 0.25] at y = 1.5 (normal -y).
    orig = np.array([[0.1, 0.0, -0.2], [0.3, 0.0, 0.0], [-0.24, 1.0, 0.24]], np.float32)
    dirs = np.tile(np.array([[0, 1.0, 0]], np.float32), (3, 1))
    h = o.trace(orig, dirs)
    assert list(h["instance"]) == [0, -1, 0]
    np.testing.assert_allclose(h["t"][[0, 2]], [1.5, 0.5], rtol=1e-6)
    np.testing.assert_allclose(h["position"][0], [0.1, 1.5, -0.2], atol=1e-6)


# ------------------------------------------------------------------ BVH --
def _box(p, tu, tv, su, sv):
    tu = tu / np.linalg.norm(tu, axis=-1, keepdims=True)
    tv = tv / np.linalg.norm(tv, axis=-1, keepdims=True)
    n = np.cross(tu, tv)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    h = np.sqrt((1.0001 * su * tu) ** 2 + (1.0001 * sv * tv) ** 2) + 0.001 * np.abs(n)
    return p - h, p + h


@pytest.mark.parametrize("n", [1, 3, 4, 5, 64, 1000, 4096])
def test_bvh_invariants(n):
    spec = scene.config_scene(0, n=n, width=8, height=8, seed=3)
    o = Oracle.from_spec(spec, dict(photons=100))
    b = o.surfel_bvh()
    pts = o.get_points()
    cnt, lf, bd = b["count"], b["left_first"], b["bounds"].astype(np.float64)
    perm = b["permutation"]
    assert sorted(perm.tolist()) == list(range(n))
    # points are the input points in BVH order
    np.testing.assert_array_equal(pts["position"], spec.points["position"][perm])
    leaves = cnt > 0
    assert cnt[leaves].sum() == n
    assert cnt.max() <= 4
    if n & (n - 1) == 0 and n >= 4:
        assert len(cnt) == 2 * (n // 4) - 1
    seen = np.zeros(n, bool)
    for i in np.nonzero(leaves)[0]:
        s = slice(lf[i], lf[i] + cnt[i])
        assert not seen[s].any()
        seen[s] = True
        lo, hi = _box(pts["position"][s].astype(np.float64), pts["tangent_u"][s], pts["tangent_v"][s],
                      pts["scale"][s, :1], pts["scale"][s, 1:])
        np.testing.assert_allclose(bd[i, :3], lo.min(0), atol=1e-6)
        np.testing.assert_allclose(bd[i, 3:], hi.max(0), atol=1e-6)
    assert seen.all()
    for i in np.nonzero(~leaves)[0]:
        for c in (lf[i], lf[i] + 1):
            assert np.all(bd[c, :3] >= bd[i, :3]) and np.all(bd[c, 3:] <= bd[i, 3:])
        # median split: the left child holds floor(count/2) of the parent's leaves
        def size(k):
            return cnt[k] if cnt[k] else size(lf[k]) + size(lf[k] + 1)
        assert size(lf[i]) == size(i) // 2


def _brute_closest(pts, o, d):
    """Float64 closest hit over all surfels (independent of the BVH)."""
    p = pts["position"].astype(np.float64)
    tu = pts["tangent_u"].astype(np.float64)
    tv = pts["tangent_v"].astype(np.float64)
    u = tu / np.linalg.norm(tu, axis=1, keepdims=True)
    v = tv - u * np.sum(u * tv, 1, keepdims=True)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    nrm = np.cross(u, v)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    nd = d @ nrm.T  # rays x surfels
    t = np.sum(nrm * p, 1)[None, :] - o @ nrm.T
    with np.errstate(divide="ignore", invalid="ignore"):
        t = t / nd
    x = o[:, None, :] + d[:, None, :] * t[..., None]
    r = x - p[None]
    uu = np.sum(r * tu[None], -1) / pts["scale"][None, :, 0]
    vv = np.sum(r * tv[None], -1) / pts["scale"][None, :, 1]
    ok = (np.abs(nd) >= 1e-6) & (t > 1e-5) & (uu * uu + vv * vv < 1.0)
    t = np.where(ok, t, np.inf)
    return np.argmin(t, 1), np.min(t, 1)


def test_closest_hit_equals_brute_force():
    spec = scene.config_scene(0, n=300, width=8, height=8, seed=5)
    spec.points["scale"] *= 3.0  # larger surfels: more overlaps along each ray
    o = Oracle.from_spec(spec, dict(photons=100))
    pts = o.get_points()
    rng = np.random.default_rng(9)
    org = rng.uniform(-1.2, 1.2, (4000, 3))
    dr = rng.normal(size=(4000, 3))
    dr /= np.linalg.norm(dr, axis=1, keepdims=True)
    h = o.trace(org.astype(np.float32), dr.astype(np.float32))
    idx, tb = _brute_closest(pts, org.astype(np.float32).astype(np.float64), dr.astype(np.float32).astype(np.float64))
    cloud = h["instance"] == 1
    hit_bf = np.isfinite(tb)
    # The light quad can be closer than any surfel; compare cloud hits only
    # where the brute force has no near-tie (|dt| < 1e-4).
    assert cloud.sum() > 300
    near = np.sort(np.stack([tb]), 0)
    sel = cloud & hit_bf
    assert np.mean(h["primitive"][sel] == idx[sel]) > 0.999
    np.testing.assert_allclose(h["t"][sel], tb[sel], rtol=1e-4)
    # every brute-force hit closer than the light is found by the BVH
    miss = hit_bf & (h["instance"] < 0)
    assert miss.sum() == 0


# ------------------------------------------------------- photons / gather --
@pytest.fixture(scope="module")
def c0_forward():
    spec = scene.config_scene(0, n=2048, width=32, height=32, seed=1)
    o = Oracle.from_spec(spec, dict(photons=20000))
    img = o.render_forward()
    return spec, o, img


def test_photon_power_closed_form(c0_forward):
    spec, o, _ = c0_forward
    ph = o.photons()
    assert ph["count"] == ph["attempted"] > 0
    # First-bounce photons carry Phi * lightCount / emitted (RayGenEmitter,
    # PrimalKernels.cpp:44-45); surfel deposits divide by qReflect = 0.5.
    phi = np.float32(spec.meshes[0].power)
    unit = np.float32(phi * np.float32(1.0)) * np.float32(1.0 / 20000)
    p = ph["power"][:, 0]
    assert np.isclose(p, unit * 2, rtol=1e-6).any() or np.isclose(p, unit, rtol=1e-6).any()
    assert np.all(ph["power"] >= 0)
    np.testing.assert_allclose(np.linalg.norm(ph["direction"], axis=1), 1.0, atol=1e-6)


def test_gather_equals_brute_force(c0_forward):
    spec, o, _ = c0_forward
    ph = o.photons()
    rng = np.random.default_rng(4)
    sel = rng.choice(ph["count"], 200, replace=False)
    q = ph["position"][sel] + rng.normal(0, 0.01, (200, 3)).astype(np.float32)
    nrm = rng.normal(size=(200, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm = nrm.astype(np.float32)
    got = o.gather(q, nrm)
    r = np.float32(spec.settings["gather_radius"])
    d2 = np.sum((ph["position"][None].astype(np.float64) - q[:, None].astype(np.float64)) ** 2, -1)
    cos = -(nrm.astype(np.float64) @ ph["direction"].astype(np.float64).T)
    w = (d2 <= float(r) ** 2) & (cos >= 0)
    want = (w @ ph["power"].astype(np.float64)) / (math.pi * float(r) ** 2)
    assert (want > 0).sum() > 50
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-9)


def test_postprocess_closed_form(c0_forward):
    _, _, img = c0_forward
    raw, ldr = img["camera1_raw"], img["camera1"]
    np.testing.assert_allclose(ldr[..., :3], np.clip(np.maximum(raw[..., :3], 0) ** (1 / 2.2), 0, 1), rtol=2e-6,
                               atol=1e-7)
    assert np.all(ldr[..., 3] == 1) and np.allclose(raw[..., 3], 1.0, atol=1e-6)


def test_single_surfel_compositing_closed_form():
    # Pixels that see only the surfel: C = alpha_geom * eta * E * albedo / pi
    # (nothing behind it), identical in every gather pass.  The pixel rays are
    # rebuilt from the seed schedule (mt19937_64(seed): emitter, 2 bounces,
    # then the camera; SyclBridge.cpp:83-143) and the per-pixel stream.
    spec = _one_surfel(opacity=0.6, su=1.5, sv=1.5)
    o = Oracle.from_spec(spec, dict(photons=20000, bounces=2, forward_passes=1, gather_passes=3, seed=7,
                                    gather_radius=0.1))
    img = o.render_forward()["camera1_raw"]
    seeds = (C.c_uint64 * 4)()
    L.orc_mt19937_64(7, 4, seeds)
    cam_seed = seeds[3] & 0xFFFFFFFF
    W = H = 16
    f = np.float32(0.5 * H / math.tan(0.5 * math.radians(45.0)))
    pix, dirs = [], []
    for py in range(6, 10):
        for px in range(6, 10):
            pi = py * W + px
            j = (C.c_float * 2)()
            L.orc_item_rng(cam_seed, pi, 2, j)
            u = px + 0.5 + (j[0] - 0.5)
            v = H - (py + 0.5 + (j[1] - 0.5))
            d = np.array([(2 * u / W - 1) * (0.5 * W) / f, (2 * v / H - 1) * (0.5 * H) / f, -1.0])
            dirs.append(d / np.linalg.norm(d))
            pix.append((py, px))
    dirs = np.array(dirs, np.float32)
    org = np.tile(np.array([[0, 0, 4.0]], np.float32), (len(dirs), 1))
    h = o.trace(org, dirs)
    assert np.all(h["instance"] == 1)
    E = o.gather(h["position"], np.tile(np.array([[0, 0, 1.0]], np.float32), (len(dirs), 1)))
    assert (E[:, 0] > 0).sum() >= 4
    want = (h["alpha"] * 0.6)[:, None] * E * 0.8 / math.pi
    got = np.array([img[py, px, :3] for py, px in pix])
    np.testing.assert_allclose(got, want, rtol=2e-4, atol=1e-9)


def test_l2_gradient_closed_form():
    # losses.py:28-45: dC/dI = (I - T) / N for C = 1/2 mean((I - T)^2).
    rng = np.random.default_rng(0)
    I = rng.random((4, 5, 3)).astype(np.float32)
    T = rng.random((4, 5, 3)).astype(np.float32)
    eps = 1e-3
    g = (I - T) / I.size
    k = (1, 2, 0)
    Ip, Im = I.astype(np.float64).copy(), I.astype(np.float64).copy()
    Ip[k] += eps
    Im[k] -= eps
    C = lambda X: 0.5 * np.mean((X - T) ** 2)
    assert (C(Ip) - C(Im)) / (2 * eps) == pytest.approx(g[k], rel=1e-6)


def test_photon_reservoir_is_unbiased():
    # depositPhotonSurface keeps a uniform sample of all attempts once the map
    # is full (reservoir sampling) and the gather scales by T / D: the
    # irradiance estimate must agree with the full map's within Monte Carlo
    # noise, and the stored photons must include late attempts.
    from differentiablepointrendering_amd import scene as sc
    from oracle.oracle import Oracle

    spec = sc.config_scene(0, n=1024, width=32, height=32)
    st = dict(photons=400_000, bounces=4, forward_passes=1, gather_passes=1, gather_radius=0.1)
    full = Oracle.from_spec(spec, st)
    full.render_forward()
    pf = full.photons()
    cap = pf["count"] // 3
    part = Oracle.from_spec(spec, dict(st, photon_map_capacity=cap))
    part.render_forward()
    pp = part.photons()
    assert pp["count"] == cap and pp["attempted"] > 2 * cap
    # The first `cap` attempts are the same in both runs (no reservoir draw
    # before them); about 1 - cap / T of the stored photons are later ones.
    early = {tuple(x) for x in pf["position"][:cap].tolist()}
    later = sum(tuple(x) not in early for x in pp["position"].tolist()) / cap
    assert abs(later - (1 - cap / pp["attempted"])) < 0.05, later
    rng = np.random.default_rng(3)
    idx = rng.choice(pf["count"], 2000, replace=False)
    q = pf["position"][idx]
    n = -pf["direction"][idx]
    ef, ep = full.gather(q, n), part.gather(q, n)
    rel = abs(ep.sum() - ef.sum()) / ef.sum()
    assert rel < 0.05, rel


def test_view_passes_are_seeded_consecutively():
    """A view of spp passes seeded s accumulates the passes seeded s, s+1, ...
    with weight 1/spp each, in pass order (f = C * (1/spp) + f)."""
    spec = scene.config_scene(0, n=512, width=24, height=24)
    o = Oracle.from_spec(spec, dict(photons=5000))
    o.build_photon_map()
    cam = scene.look_at_camera("v", (1.0, 1.0, 3.7), (0, 0, 0), (0, 1, 0), 20, 16, 45.0)
    a = o.render_view(cam, 1, 41)["raw"]
    b = o.render_view(cam, 1, 42)["raw"]
    ab = o.render_view(cam, 2, 41)["raw"]
    half = np.float32(0.5)
    want = b * half + a * half
    assert np.array_equal(ab.view(np.uint32), want.view(np.uint32))
    assert not np.array_equal(a, b)
    # The photon map of build_photon_map is the one render_forward builds.
    p0 = o.photons()
    o.render_forward()
    p1 = o.photons()
    assert np.array_equal(p0["position"], p1["position"]) and p0["count"] == p1["count"] > 0


def test_sample_offsets_partition_the_image():
    """The CPU baseline's bounded sample: handles with pixel stride k and
    offsets 0..k-1 render disjoint pixel sets whose images add up to the full
    render (every pixel is computed independently of the others over the
    same photon map)."""
    spec = scene.config_scene(0, n=512, width=24, height=20)
    st = dict(photons=4000, gather_passes=2)
    full = Oracle.from_spec(spec, st)
    ref = full.render_forward()[spec.cameras[0].name + "_raw"]
    full.close()
    acc = np.zeros_like(ref)
    seen = np.zeros(ref.shape[:2], np.int32)
    for off in range(3):
        o = Oracle.from_spec(spec, st)
        o.set_sample(1, 3, off)
        img = o.render_forward()[spec.cameras[0].name + "_raw"]
        o.close()
        seen += img[..., 3] > 0
        acc += img
    assert np.array_equal(seen, np.ones_like(seen))
    assert np.array_equal(acc, ref)


def test_oracle_thread_count_does_not_change_results():
    """The oracle spreads closest-hit queries, camera pixels and event gathers
    over threads; everything order-dependent stays sequential, so one thread
    and many give identical photon maps, images and gradients."""
    spec = scene.config_scene(0, n=2048, width=24, height=24, seed=2)
    st = dict(photons=8000, bounces=4, gather_passes=2, adjoint_passes=2, adjoint_bounces=4, rr_start=2,
              full_gradients=1)
    out = []
    for nt in (1, 4):
        L.orc_set_threads(nt)
        o = Oracle.from_spec(spec, st)
        img = o.render_forward()["camera1_raw"]
        dl = (img[..., :3] / img[..., :3].size).astype(np.float32)
        g = o.render_backward({"camera1": dl})
        out.append((img, o.photons(), g, o.full_gradients(), o.stats()))
        o.close()
    L.orc_set_threads(0)
    (a, pa, ga, fa, sa), (b, pb, gb, fb, sb) = out
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(pa["position"], pb["position"]) and np.array_equal(pa["power"], pb["power"])
    assert np.array_equal(ga.view(np.uint32), gb.view(np.uint32))
    for k in fa:
        assert np.array_equal(fa[k].view(np.uint32), fb[k].view(np.uint32)), k
    for k in ("photon_queries", "gather_queries", "adjoint_queries", "gradient_events", "work_photon",
              "work_gather", "work_adjoint"):
        assert sa[k] == sb[k], k
    assert np.count_nonzero(ga) > 0


# ------------------------------------------------- adjoint seed schedule --
def test_adjoint_seed_schedule():
    """adjoint_seed is the adjoint's base seed (Random::number, default 42):
    the default equals 42 given explicitly; with adjoint_decorrelate the
    2-sample adjoint from seed s is the mean of the 1-sample adjoints from
    s and s + 1 (sample k seeds from s + k); without it both samples repeat
    the seed-s sample."""
    spec = scene.config_scene(0, n=4096, width=32, height=32)
    st = dict(photons=20_000, gather_passes=1, full_gradients=0)

    def grad(**kw):
        o = Oracle.from_spec(spec, dict(st, **kw))
        raw = o.render_forward()["camera1_raw"]
        dl = (raw[..., :3] / raw[..., :3].size).astype(np.float32)
        g = o.render_backward({"camera1": dl})
        o.close()
        return g

    base = grad(adjoint_passes=1)
    assert np.abs(base).max() > 0
    assert np.array_equal(base, grad(adjoint_passes=1, adjoint_seed=42))
    s = 1000
    one_a = grad(adjoint_passes=1, adjoint_seed=s)
    one_b = grad(adjoint_passes=1, adjoint_seed=s + 1)
    assert not np.array_equal(one_a, one_b)
    two = grad(adjoint_passes=2, adjoint_seed=s, adjoint_decorrelate=1)
    scale = np.abs(two).max()
    np.testing.assert_allclose(two, 0.5 * (one_a + one_b), rtol=1e-5, atol=1e-6 * scale)
    rep = grad(adjoint_passes=2, adjoint_seed=s)
    np.testing.assert_allclose(rep, one_a, rtol=1e-5, atol=1e-6 * scale)


# --------------------------------------------------------- debug images --
def test_debug_opacity_image_sums_to_the_opacity_gradient():
    """debug_images: every adjoint event adds its opacity-gradient
    contribution to its pixel's R, G and B (atomicAddFloat4ToImage,
    KernelHelpers.h:1489-1502) and to its surfel's gradient, so the image
    sums to the gradient's sum; alpha is 1 exactly where events landed, and
    the setting changes no gradient."""
    spec = scene.config_scene(0, n=4096, width=32, height=32)
    st = dict(photons=20_000, gather_passes=1, adjoint_passes=2, adjoint_bounces=3)

    def run(debug):
        o = Oracle.from_spec(spec, dict(st, debug_images=debug))
        raw = o.render_forward()["camera1_raw"]
        dl = (raw[..., :3] / raw[..., :3].size).astype(np.float32)
        g = o.render_backward({"camera1": dl})
        img = o.debug_image("camera1", 32, 32) if debug else None
        o.close()
        return g, img

    g0, _ = run(0)
    g, img = run(1)
    assert np.array_equal(g, g0)
    assert np.abs(g).sum() > 0
    assert np.array_equal(img[..., 0], img[..., 1]) and np.array_equal(img[..., 0], img[..., 2])
    assert set(np.unique(img[..., 3])) <= {0.0, 1.0}
    assert np.all(img[img[..., 3] == 0.0][:, :3] == 0.0)
    assert (img[..., 3] == 1.0).sum() > 10
    np.testing.assert_allclose(img[..., 0].astype(np.float64).sum(), g.astype(np.float64).sum(), rtol=1e-5,
                               atol=1e-6 * np.abs(g).sum())
