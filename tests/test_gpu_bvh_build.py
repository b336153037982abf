"""The surfel BVH built on the GPU (setting bvh_build = 1) against the host
build (bvh_build = 2), bit for bit: the reference nodes (bounds, left_first,
count), the permutation, the point parameters in BVH order and, where a
camera is given, the rendered images (which read the device's surfel
records and wide nodes).  Sizes sit around the kernels' thresholds: 5
surfels (one split), 64 / 65 (the one-thread subtree limit), 257 and 4097
(the workgroup selection's 256-element tail), 100,003 and the plain
workload's 1,048,576; keys with many ties and duplicated positions; a
training-style loop (apply_point_optimization, rebuild_bvh, parameters
fetched again); point edits that rebuild; and a scene with a -0 box
coordinate, which both settings build on the host.  The host build follows
std::nth_element (BVH.ixx:49-133, SceneBuild.cpp:642-772) and is pinned to
the oracle in test_gpu_parity.py and test_gpu_configs.py.
"""
import numpy as np
import pytest

from differentiablepointrendering_amd import scene

pytestmark = pytest.mark.gpu

SMALL = dict(photons=20_000, bounces=3, forward_passes=1, gather_passes=1, adjoint_bounces=2, adjoint_passes=1,
             gather_radius=0.05)


def _points(n, mode, seed=0):
    pts, _ = scene.clustered_surfels(n, seed=seed)
    pts = {k: np.ascontiguousarray(v, dtype=np.float32).copy() for k, v in pts.items()}
    rng = np.random.default_rng(seed + 11)
    if mode == "ties":
        pts["position"] = (np.round(pts["position"] * 4.0) / 4.0).astype(np.float32)
    elif mode == "duplicates":
        src = rng.integers(0, n, size=n // 3)
        pts["position"][rng.integers(0, n, size=n // 3)] = pts["position"][src]
    return pts


def _renderers(points, cameras=(), settings=None):
    from differentiablepointrendering_amd.pale import Renderer

    s = dict(SMALL, **(settings or {}))
    meshes = [scene.light_quad()]
    n = len(points["position"])
    print(f"case: bvh build n={n} cameras={len(cameras)} settings={s}", flush=True)
    dev = Renderer.from_scene(meshes, list(cameras), points, dict(s, bvh_build=1))
    host = Renderer.from_scene(meshes, list(cameras), points, dict(s, bvh_build=2))
    return dev, host


def _same(dev, host, images=False):
    assert dev.stats()["bvh_on_device"] == 1
    assert host.stats()["bvh_on_device"] == 0
    a, b = dev.surfel_bvh(), host.surfel_bvh()
    for k in ("left_first", "count", "permutation"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["bounds"].view(np.uint32), b["bounds"].view(np.uint32))
    pa, pb = dev.get_point_parameters(), host.get_point_parameters()
    for k in pb:
        assert np.array_equal(pa[k].view(np.uint32), pb[k].view(np.uint32)), k
    if images:
        ia, ib = dev.render_forward(), host.render_forward()
        for k in ib:
            assert np.array_equal(ia[k].view(np.uint32), ib[k].view(np.uint32)), k


@pytest.mark.parametrize("n", [5, 64, 65, 257, 4097, 100_003])
@pytest.mark.parametrize("mode", ["clustered", "ties", "duplicates"])
def test_device_build_equals_host_build(n, mode):
    dev, host = _renderers(_points(n, mode))
    try:
        _same(dev, host)
    finally:
        dev.close()
        host.close()


def test_device_build_plain_workload_images():
    """1,048,576 clustered surfels: the default setting (GPU from 65,536
    surfels) against the host build, BVH and a rendered image."""
    spec = scene.config_scene(1)
    cam = scene.default_camera(128, 128)
    from differentiablepointrendering_amd.pale import Renderer

    meshes = spec.meshes
    print("case: bvh build plain workload 1048576 surfels", flush=True)
    dev = Renderer.from_scene(meshes, [cam], spec.points, dict(SMALL))
    host = Renderer.from_scene(meshes, [cam], spec.points, dict(SMALL, bvh_build=2))
    try:
        _same(dev, host, images=True)
        print(f"bvh build ms: device {dev.stats()['bvh_build_ms']:.1f} host {host.stats()['bvh_build_ms']:.1f}",
              flush=True)
    finally:
        dev.close()
        host.close()


def test_device_rebuild_in_training_loop():
    """training.py:621-750's point update: apply_point_optimization with
    moved parameters, rebuild_bvh, parameters fetched again, twice; images
    and BVH equal to the host-built renderer doing the same."""
    pts = _points(80_000, "clustered", seed=3)
    cam = scene.default_camera(64, 64)
    dev, host = _renderers(pts, [cam])
    try:
        _same(dev, host, images=True)
        rng = np.random.default_rng(5)
        for _ in range(2):
            p = dev.get_point_parameters()
            step = {k: v for k, v in p.items() if k != "shape"}
            step["position"] = (step["position"] + rng.normal(0, 0.01, step["position"].shape)).astype(np.float32)
            step["opacity"] = np.clip(step["opacity"] + 0.05, 0, 1).astype(np.float32)
            for r in (dev, host):
                r.apply_point_optimization(step)
                r.rebuild_bvh()
            _same(dev, host, images=True)
    finally:
        dev.close()
        host.close()


def test_device_build_after_point_edits():
    pts = _points(70_000, "clustered", seed=4)
    dev, host = _renderers(pts)
    try:
        extra = {k: v[:1000].copy() for k, v in _points(1000, "clustered", seed=9).items()}
        for r in (dev, host):
            r.add_points({"new": extra})
        _same(dev, host)
        gone = {"indices": np.arange(0, 5000, 3, dtype=np.int64)}
        for r in (dev, host):
            r.remove_points(gone)
        _same(dev, host)
    finally:
        dev.close()
        host.close()


def test_negative_zero_box_is_built_on_the_host():
    """A box coordinate of -0 (a zero-size surfel at x = -0 whose normal has
    no x part) would make min / max depend on the order they are taken in:
    bvh_build = 1 builds such a scene on the host."""
    from differentiablepointrendering_amd.pale import Renderer

    pts = _points(70_000, "clustered", seed=6)
    pts["position"][123, 0] = np.float32(-0.0)
    pts["tangent_u"][123] = (1.0, 0.0, 0.0)
    pts["tangent_v"][123] = (0.0, 1.0, 0.0)
    pts["scale"][123] = (0.0, 0.0)
    print("case: bvh build with a -0 box coordinate", flush=True)
    dev = Renderer.from_scene([scene.light_quad()], [], pts, dict(SMALL, bvh_build=1))
    host = Renderer.from_scene([scene.light_quad()], [], pts, dict(SMALL, bvh_build=2))
    try:
        assert dev.stats()["bvh_on_device"] == 0
        a, b = dev.surfel_bvh(), host.surfel_bvh()
        for k in ("left_first", "count", "permutation"):
            assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(a["bounds"].view(np.uint32), b["bounds"].view(np.uint32))
    finally:
        dev.close()
        host.close()
