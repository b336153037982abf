"""Drop-in for the reference's pybind11 module ``pale`` (Pale/pale_bindings.cpp).

``Renderer`` has the reference's constructor and methods with the same
argument names, return shapes and errors (RuntimeError for bad shape / dtype
/ key / size, IndexError for out-of-range indices), backed by the HIP library
through the C-ABI in include/pale_abi.h.  The reference drivers
(python/training.py, render_hooks.py, render_adjoint.py, finite_difference/*)
can use it by making it importable as ``pale``::

    import sys, differentiablepointrendering_amd.pale as pale
    sys.modules["pale"] = pale

Compatibility extensions (supersets of the reference):
  * render_backward accepts a bare ndarray (every adjoint camera of that
    size) and (H,W,4) images (alpha ignored);
  * the gradient dict also carries "color" (alias of "albedo");
  * render_forward returns a dict whose ``np.asarray`` is the first camera's
    LDR RGB image;
  * settings key "ldr_channels" (default 4, the reference's (H,W,4) LDR):
    3 makes render_forward's LDR images (H,W,3), the shape training.py
    compares with its (H,W,3) targets (io_utils.read_rgb_exr); the "_raw"
    HDR images stay (H,W,4) either way.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _abi
from ._lib import check, load

_fp = C.POINTER(C.c_float)
# Keys of images["debug"][camera] in the reference's order
# (pale_bindings.cpp:646-715).
DEBUG_IMAGE_KEYS = ("position_x", "position_y", "position_z", "rotation", "scale", "albedo", "opacity", "beta")


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_fp)


class ForwardImages(dict):
    """render_forward result: {name: (H,W,4) LDR (or (H,W,3) with
    ldr_channels=3), name + "_raw": (H,W,4) HDR}."""

    def __array__(self, dtype=None, copy=None):
        for k, v in self.items():
            if not k.endswith("_raw"):
                a = v[..., :3]
                return a.astype(dtype) if dtype is not None else a
        return np.zeros((0, 0, 3), np.float32)


class Renderer:
    """pale.Renderer (pale_bindings.cpp:59-1650) on the MI355X."""

    def __init__(self, assetRootDir: str, sceneXml: str = "cbox_custom.xml", pointCloudFile: str = "initial.ply",
                 settings: Optional[dict] = None):
        from . import assets

        # The reference changes the working directory to the asset root
        # (pale_bindings.cpp:66); paths are resolved against it here instead.
        meshes, cameras, points = assets.load_scene(assetRootDir, sceneXml, pointCloudFile)
        self._init(meshes, cameras, points, settings or {})

    @classmethod
    def from_scene(cls, meshes, cameras, points: Dict[str, np.ndarray], settings: Optional[dict] = None):
        """Renderer over an in-memory scene (meshes/cameras of ``_abi``)."""
        self = cls.__new__(cls)
        self._init(meshes, cameras, points, settings or {})
        return self

    @classmethod
    def from_spec(cls, spec, settings: Optional[dict] = None):
        s = dict(spec.settings)
        s.update(settings or {})
        return cls.from_scene(spec.meshes, spec.cameras, spec.points, s)

    def _init(self, meshes, cameras, points, settings):
        self._L = load()
        self._settings_dict = dict(settings)
        ldr_channels = int(settings.get("ldr_channels", 4))
        if ldr_channels not in (3, 4):
            raise RuntimeError(f"ldr_channels must be 3 or 4, got {ldr_channels}")
        self._ldr_channels = ldr_channels
        holder = _abi.SceneHolder(list(meshes), list(cameras), points)
        self._settings = _abi.make_settings(settings)
        h = C.c_void_p()
        check(self._L.pale_create(C.byref(holder.desc), C.byref(self._settings), C.byref(h)))
        self._h = h
        self._cams = []
        n = C.c_uint32()
        check(self._L.pale_camera_count(self._h, C.byref(n)))
        for i in range(n.value):
            name = C.create_string_buffer(_abi.PALE_NAME_LEN)
            w, hh, adj = C.c_uint32(), C.c_uint32(), C.c_int32()
            check(self._L.pale_camera_info(self._h, i, name, C.byref(w), C.byref(hh), C.byref(adj)))
            self._cams.append((name.value.decode(), w.value, hh.value, bool(adj.value)))

    @property
    def full_gradients(self) -> bool:
        return bool(self._settings.full_gradients)

    def close(self):
        if getattr(self, "_h", None):
            self._L.pale_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ cameras --
    def get_camera_names(self) -> List[str]:
        return [c[0] for c in self._cams]

    # ------------------------------------------------------------ forward --
    def render_forward(self, camera_name: str = "") -> ForwardImages:
        cams = [c for c in self._cams if camera_name == "" or c[0] == camera_name]
        ldr = [np.empty((c[2], c[1], 4), np.float32) for c in cams]
        raw = [np.empty((c[2], c[1], 4), np.float32) for c in cams]
        k = max(1, len(cams))
        lp = (_fp * k)(*[_p(a) for a in ldr])
        rp = (_fp * k)(*[_p(a) for a in raw])
        n = C.c_uint32()
        check(self._L.pale_render_forward(self._h, camera_name.encode(), lp, rp, C.byref(n)))
        out = ForwardImages()
        for c, a, b in zip(cams, ldr, raw):
            out[c[0]] = a if self._ldr_channels == 4 else np.ascontiguousarray(a[..., :3])
            out[c[0] + "_raw"] = b
        return out

    # ----------------------------------------------------------- backward --
    def _adjoint_inputs(self, target) -> Tuple[List[str], List[np.ndarray]]:
        adj = [c for c in self._cams if c[3]]
        if isinstance(target, np.ndarray):
            target = {c[0]: target for c in adj if target.ndim == 3 and target.shape[:2] == (c[2], c[1])}
        names, arrs = [], []
        for name, w, h, _ in adj:
            if name not in target:
                continue
            a = np.asarray(target[name])
            if a.ndim != 3 or a.shape[2] not in (3, 4):
                raise RuntimeError(f"render_backward: target image for camera '{name}' must be HxWx3 float32")
            if a.dtype.itemsize != 4 or a.dtype.kind != "f":
                raise RuntimeError(f"render_backward: target image for camera '{name}' must have dtype float32")
            if a.shape[0] != h or a.shape[1] != w:
                raise RuntimeError(f"render_backward: resolution mismatch for camera '{name}': target image is "
                                   f"{a.shape[1]}x{a.shape[0]}, but sensor is {w}x{h}")
            names.append(name)
            arrs.append(np.ascontiguousarray(a[..., :3], dtype=np.float32))
        return names, arrs

    def render_backward(self, targetRgb32f) -> Tuple[dict, dict]:
        names, arrs = self._adjoint_inputs(targetRgb32f)
        n = self.point_count
        grads = {k: np.empty((n, w) if w else (n,), np.float32) for k, w in _abi.POINT_SHAPES.items()}
        gp = (_fp * 8)(*[_p(grads[k]) for k in _abi.POINT_KEYS])
        srcs = [np.empty((a.shape[0], a.shape[1], 4), np.float32) for a in arrs]
        k = max(1, len(names))
        cn = (C.c_char_p * k)(*[s.encode() for s in names])
        ap = (_fp * k)(*[_p(a) for a in arrs])
        sp = (_fp * k)(*[_p(a) for a in srcs])
        check(self._L.pale_render_backward(self._h, len(names), cn, ap, gp, sp))
        grads["color"] = grads["albedo"]
        images = {"adjoint_source": {nm: s for nm, s in zip(names, srcs)}}
        if self._settings.debug_images:
            # images["debug"] (pale_bindings.cpp:627-726): per camera the eight
            # gradient images; the reference's kernels write only "opacity".
            debug = {}
            for nm, a in zip(names, arrs):
                per = {}
                for key in DEBUG_IMAGE_KEYS:
                    per[key] = np.zeros((a.shape[0], a.shape[1], 4), np.float32)
                check(self._L.pale_get_debug_image(self._h, nm.encode(), _p(per["opacity"])))
                debug[nm] = per
            images["debug"] = debug
        return grads, images

    # ------------------------------------------------ device (torch) tensors --
    @staticmethod
    def _dptr(t):
        return C.cast(C.c_void_p(t.data_ptr()), _fp)

    def render_forward_device(self, camera_name: str = "") -> dict:
        """render_forward into torch tensors on the renderer's GPU:
        {name: (H,W,4) LDR, name + "_raw": (H,W,4) HDR}."""
        import torch

        cams = [c for c in self._cams if camera_name == "" or c[0] == camera_name]
        ldr = [torch.empty((c[2], c[1], 4), dtype=torch.float32, device="cuda") for c in cams]
        raw = [torch.empty((c[2], c[1], 4), dtype=torch.float32, device="cuda") for c in cams]
        k = max(1, len(cams))
        lp = (_fp * k)(*[self._dptr(a) for a in ldr])
        rp = (_fp * k)(*[self._dptr(a) for a in raw])
        n = C.c_uint32()
        torch.cuda.current_stream().synchronize()
        check(self._L.pale_render_forward_device(self._h, camera_name.encode(), lp, rp, C.byref(n)))
        out = {}
        for c, a, b in zip(cams, ldr, raw):
            out[c[0]] = a if self._ldr_channels == 4 else a[..., :3].contiguous()
            out[c[0] + "_raw"] = b
        return out

    def render_backward_device(self, dloss: dict) -> dict:
        """render_backward from (H,W,3) dL/dI torch tensors on the GPU; returns
        the gradient dict as torch tensors on the GPU."""
        import torch

        names = [c[0] for c in self._cams if c[3] and c[0] in dloss]
        imgs = []
        for nm in names:
            cam = next(c for c in self._cams if c[0] == nm)
            t = dloss[nm]
            if t.dim() != 3 or t.shape[2] not in (3, 4):
                raise RuntimeError("adjoint image for camera '" + nm + "' must be (H,W,3) or (H,W,4) float32")
            if t.shape[0] != cam[2] or t.shape[1] != cam[1]:
                raise RuntimeError(f"Adjoint image size mismatch for camera '{nm}': expected {cam[2]}x{cam[1]}, "
                                   f"got {t.shape[0]}x{t.shape[1]}")
            imgs.append(t[..., :3].to(device="cuda", dtype=torch.float32).contiguous())
        n = self.point_count
        grads = {k: torch.empty((n, w) if w else (n,), dtype=torch.float32, device="cuda")
                 for k, w in _abi.POINT_SHAPES.items()}
        gp = (_fp * 8)(*[self._dptr(grads[k]) for k in _abi.POINT_KEYS])
        k = max(1, len(names))
        cn = (C.c_char_p * k)(*[s.encode() for s in names])
        ap = (_fp * k)(*[self._dptr(a) for a in imgs])
        torch.cuda.current_stream().synchronize()
        check(self._L.pale_render_backward_device(self._h, len(names), cn, ap, gp))
        grads["color"] = grads["albedo"]
        return grads

    def set_opacity_device(self, opacity) -> None:
        """Opacity of every surfel (BVH order) from a GPU tensor, without a
        BVH rebuild."""
        import torch

        t = opacity.detach().to(device="cuda", dtype=torch.float32).contiguous().reshape(-1)
        torch.cuda.current_stream().synchronize()
        check(self._L.pale_set_opacity_device(self._h, self._dptr(t), t.numel()))

    def apply_point_optimization_device(self, parameters: dict) -> None:
        """apply_point_optimization from torch tensors on the GPU (keys
        position, tangent_u, tangent_v, scale, albedo, opacity, beta; BVH
        order; same N): the parameters never leave device memory and the BVH
        is kept."""
        import torch

        n = self.point_count
        shapes = {"position": (n, 3), "tangent_u": (n, 3), "tangent_v": (n, 3), "scale": (n, 2),
                  "albedo": (n, 3), "opacity": (n,), "beta": (n,)}
        ts = []
        for k, shp in shapes.items():
            if k not in parameters:
                raise RuntimeError(f"Missing key '{k}' in parameters dict")
            t = parameters[k].detach()
            if t.numel() != math.prod(shp):
                raise RuntimeError("apply_point_optimization expects consistent point count; use densification "
                                   "API for adding/removing points.")
            ts.append(t.to(device="cuda", dtype=torch.float32).contiguous().reshape(shp))
        torch.cuda.current_stream().synchronize()
        check(self._L.pale_apply_points_device(self._h, n, *[C.c_void_p(t.data_ptr()) for t in ts]))

    # ------------------------------------------------------- served views --
    def build_photon_map(self) -> None:
        """Photon stage of render_forward only; views then gather from it."""
        check(self._L.pale_build_photon_map(self._h))

    @staticmethod
    def device_bytes() -> int:
        """Device bytes held by every renderer of the process."""
        v = C.c_uint64()
        check(load().pale_device_bytes(C.byref(v)))
        return v.value

    def view_bytes(self, width: int, height: int, spp: int, adjoint: bool) -> int:
        """Device bytes a width x height view of spp passes (with its
        adjoint) would add."""
        v = C.c_uint64()
        check(self._L.pale_view_bytes(self._h, int(width), int(height), int(spp), 1 if adjoint else 0,
                                      C.byref(v)))
        return v.value

    def views_bytes(self, sizes, spp, adjoint) -> int:
        """Device bytes render_views of views of `sizes` ((w, h) each) with
        `spp`, then render_view_backward of those with adjoint[i], would add."""
        n = len(sizes)
        w = np.ascontiguousarray([s[0] for s in sizes], dtype=np.uint32)
        h = np.ascontiguousarray([s[1] for s in sizes], dtype=np.uint32)
        sp = np.ascontiguousarray(spp, dtype=np.uint32)
        ad = np.ascontiguousarray([1 if x else 0 for x in adjoint], dtype=np.int32)
        v = C.c_uint64()
        u32 = C.POINTER(C.c_uint32)
        check(self._L.pale_views_bytes(self._h, n, w.ctypes.data_as(u32), h.ctypes.data_as(u32),
                                       sp.ctypes.data_as(u32), ad.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(v)))
        return v.value

    def view_buffer_bytes(self) -> int:
        """Device bytes of the image-sized buffers (what release_view_buffers frees)."""
        v = C.c_uint64()
        check(self._L.pale_view_buffer_bytes(self._h, C.byref(v)))
        return v.value

    def release_view_buffers(self) -> None:
        """Frees the image-sized buffers; scene and photon map stay."""
        check(self._L.pale_release_view_buffers(self._h))

    def render_view(self, camera: _abi.Camera, spp: int, seed: int) -> dict:
        """One view of the resident scene over the current photon map:
        {"ldr": (H,W,4), "raw": (H,W,4)}.  Pass k draws the seed seed + k, so
        the image depends only on the scene, the settings and (camera, spp,
        seed)."""
        d = _abi.fill_camera_desc(_abi.CameraDesc(), camera)
        ldr = np.empty((camera.height, camera.width, 4), np.float32)
        raw = np.empty((camera.height, camera.width, 4), np.float32)
        check(self._L.pale_render_view(self._h, C.byref(d), int(spp), int(seed) & 0xFFFFFFFF, _p(ldr), _p(raw)))
        return {"ldr": ldr, "raw": raw}

    def set_batch_tuning(self, slots_per_batch: int = 0, entries_per_slot: int = 0, graph_rounds: int = 0,
                         grid_fraction: float = 0.0, walk_lds_entries: int = 0) -> None:
        """Launch shape of render_views (0 = default); never changes results."""
        t = _abi.BatchTuning(int(slots_per_batch), int(entries_per_slot), int(graph_rounds), float(grid_fraction),
                             int(walk_lds_entries))
        check(self._L.pale_set_batch_tuning(self._h, C.byref(t)))

    def batch_tuning(self) -> dict:
        t = _abi.BatchTuning()
        check(self._L.pale_get_batch_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def render_views(self, cameras, spp, seeds) -> List[dict]:
        """Several views over the current photon map, batched: a list of
        {"ldr", "raw"} equal to render_view(cameras[i], spp[i], seeds[i])."""
        n = len(cameras)
        if len(spp) != n or len(seeds) != n:
            raise RuntimeError("render_views: cameras, spp and seeds must have the same length")
        descs = (_abi.CameraDesc * max(1, n))()
        for i, c in enumerate(cameras):
            _abi.fill_camera_desc(descs[i], c)
        sp = np.ascontiguousarray(spp, dtype=np.uint32)
        sd = np.ascontiguousarray(np.asarray(seeds, dtype=np.int64) & 0xFFFFFFFF, dtype=np.uint32)
        ldr = [np.empty((c.height, c.width, 4), np.float32) for c in cameras]
        raw = [np.empty((c.height, c.width, 4), np.float32) for c in cameras]
        k = max(1, n)
        lp = (_fp * k)(*[_p(a) for a in ldr])
        rp = (_fp * k)(*[_p(a) for a in raw])
        nb = C.c_uint32()
        u32 = C.POINTER(C.c_uint32)
        check(self._L.pale_render_views(self._h, n, descs, sp.ctypes.data_as(u32), sd.ctypes.data_as(u32), lp, rp,
                                        C.byref(nb)))
        self.last_batches = nb.value
        return [{"ldr": a, "raw": b} for a, b in zip(ldr, raw)]

    def render_view_backward_device(self, camera: _abi.Camera, spp: int, dloss) -> dict:
        """render_view_backward from an (H,W,3) dL/dI torch tensor on the GPU;
        the gradient dict as torch tensors on the GPU."""
        import torch

        if dloss.dim() != 3 or dloss.shape[0] != camera.height or dloss.shape[1] != camera.width:
            raise RuntimeError(f"render_view_backward: dL/dI must be ({camera.height},{camera.width},3)")
        dl = dloss[..., :3].to(device="cuda", dtype=torch.float32).contiguous()
        d = _abi.fill_camera_desc(_abi.CameraDesc(), camera)
        n = self.point_count
        grads = {k: torch.empty((n, w) if w else (n,), dtype=torch.float32, device="cuda")
                 for k, w in _abi.POINT_SHAPES.items()}
        gp = (_fp * 8)(*[self._dptr(grads[k]) for k in _abi.POINT_KEYS])
        torch.cuda.current_stream().synchronize()
        check(self._L.pale_render_view_backward_device(self._h, C.byref(d), int(spp), self._dptr(dl), gp))
        grads["color"] = grads["albedo"]
        return grads

    def render_views_backward(self, cameras, spp, dloss) -> dict:
        """Adjoint of several views (dL/dI (H,W,3) each) over the current
        photon map; the gradient dict of render_backward, summed over views."""
        n = len(cameras)
        if len(spp) != n or len(dloss) != n:
            raise RuntimeError("render_views_backward: cameras, spp and dloss must have the same length")
        descs = (_abi.CameraDesc * max(1, n))()
        arrs = []
        for i, c in enumerate(cameras):
            _abi.fill_camera_desc(descs[i], c)
            a = np.ascontiguousarray(np.asarray(dloss[i])[..., :3], dtype=np.float32)
            if a.shape != (c.height, c.width, 3):
                raise RuntimeError(f"render_views_backward: dL/dI {i} must be ({c.height},{c.width},3), got {a.shape}")
            arrs.append(a)
        sp = np.ascontiguousarray(spp, dtype=np.uint32)
        pts = self.point_count
        grads = {k: np.empty((pts, w) if w else (pts,), np.float32) for k, w in _abi.POINT_SHAPES.items()}
        gp = (_fp * 8)(*[_p(grads[k]) for k in _abi.POINT_KEYS])
        ap = (_fp * max(1, n))(*[_p(a) for a in arrs])
        nb = C.c_uint32()
        check(self._L.pale_render_views_backward(self._h, n, descs, sp.ctypes.data_as(C.POINTER(C.c_uint32)), ap,
                                                 gp, C.byref(nb)))
        self.last_batches = nb.value
        grads["color"] = grads["albedo"]
        return grads

    def render_view_device(self, camera: _abi.Camera, spp: int, seed: int) -> dict:
        """render_view into torch tensors on the renderer's GPU."""
        import torch

        d = _abi.fill_camera_desc(_abi.CameraDesc(), camera)
        ldr = torch.empty((camera.height, camera.width, 4), dtype=torch.float32, device="cuda")
        raw = torch.empty((camera.height, camera.width, 4), dtype=torch.float32, device="cuda")
        torch.cuda.current_stream().synchronize()
        check(self._L.pale_render_view_device(self._h, C.byref(d), int(spp), int(seed) & 0xFFFFFFFF,
                                              C.c_void_p(ldr.data_ptr()), C.c_void_p(raw.data_ptr())))
        return {"ldr": ldr, "raw": raw}

    def render_view_backward_noalloc(self, camera: _abi.Camera, spp: int, dloss, opacity_grad) -> None:
        """Opacity gradient of render_view_backward into the caller's device
        tensor `opacity_grad` ((N,) float32), from an (H,W,3) float32 dL/dI
        device tensor, with no device allocation anywhere on the way (the
        library's static scratch array holds the intermediates; views of at
        most 256 x 256 pixels)."""
        n = self.point_count
        if tuple(dloss.shape) != (camera.height, camera.width, 3) or not dloss.is_contiguous():
            raise RuntimeError(f"render_view_backward_noalloc: dL/dI must be a contiguous ({camera.height},"
                               f"{camera.width},3) tensor")
        if tuple(opacity_grad.shape) != (n,) or not opacity_grad.is_contiguous():
            raise RuntimeError(f"render_view_backward_noalloc: the gradient must be a contiguous ({n},) tensor")
        if not (dloss.is_cuda and opacity_grad.is_cuda):
            raise RuntimeError("render_view_backward_noalloc: tensors must be on the GPU")
        d = _abi.fill_camera_desc(_abi.CameraDesc(), camera)
        import torch

        torch.cuda.current_stream().synchronize()
        check(self._L.pale_render_view_backward_noalloc(self._h, C.byref(d), int(spp), C.c_void_p(dloss.data_ptr()),
                                                        C.c_void_p(opacity_grad.data_ptr())))

    def render_view_backward(self, camera: _abi.Camera, spp: int, dloss_rgb: np.ndarray) -> dict:
        """Adjoint of one view: spp adjoint samples of dL/dI (H,W,3) through
        the current photon map; the gradient dict of render_backward."""
        a = np.ascontiguousarray(np.asarray(dloss_rgb)[..., :3], dtype=np.float32)
        if a.shape != (camera.height, camera.width, 3):
            raise RuntimeError(f"render_view_backward: dL/dI must be ({camera.height},{camera.width},3), "
                               f"got {a.shape}")
        d = _abi.fill_camera_desc(_abi.CameraDesc(), camera)
        n = self.point_count
        grads = {k: np.empty((n, w) if w else (n,), np.float32) for k, w in _abi.POINT_SHAPES.items()}
        gp = (_fp * 8)(*[_p(grads[k]) for k in _abi.POINT_KEYS])
        check(self._L.pale_render_view_backward(self._h, C.byref(d), int(spp), _p(a), gp))
        grads["color"] = grads["albedo"]
        return grads

    # ------------------------------------------------------------- points --
    @property
    def point_count(self) -> int:
        n = C.c_uint32()
        check(self._L.pale_point_count(self._h, C.byref(n)))
        return n.value

    def get_point_parameters(self) -> Dict[str, np.ndarray]:
        n = self.point_count
        out = {k: np.empty((n, w) if w else (n,), np.float32) for k, w in _abi.POINT_SHAPES.items()}
        check(self._L.pale_get_points(self._h, *[_p(out[k]) for k in _abi.POINT_KEYS]))
        return out

    @staticmethod
    def _points_from_dict(d: dict, n: Optional[int], ctx: str) -> _abi.PointsHolder:
        for key in ("position", "tangent_u", "tangent_v", "scale", "albedo", "opacity", "beta"):
            if key not in d:
                raise RuntimeError(f"{ctx}: missing key '{key}'")
        arrays = {}
        for key in ("position", "tangent_u", "tangent_v", "scale", "albedo", "opacity", "beta"):
            a = np.asarray(d[key])
            w = _abi.POINT_SHAPES[key]
            m = a.shape[0] if a.ndim >= 1 else -1
            want = (m, w) if w else (m,)
            if a.shape != want or (n is not None and m != n):
                shape = f"(N,{w})" if w else "(N,)"
                raise RuntimeError(f"Expected '{key}' to have shape {shape}")
            if a.dtype.itemsize != 4 or a.dtype.kind != "f":
                raise RuntimeError(f"Expected '{key}' to be float32")
            arrays[key] = a
        return _abi.PointsHolder(arrays)

    def apply_point_optimization(self, parameters: dict) -> None:
        if "position" not in parameters:
            return
        pos = np.asarray(parameters["position"])
        if pos.ndim != 2 or pos.shape[1] != 3:
            raise RuntimeError("Expected 'position' to have shape (N,3)")
        if pos.shape[0] == 0:
            return
        if pos.shape[0] != self.point_count:
            raise RuntimeError("apply_point_optimization expects consistent point count; use densification API "
                               "for adding/removing points.")
        holder = self._points_from_dict(parameters, pos.shape[0], "apply_point_optimization")
        check(self._L.pale_apply_points(self._h, C.byref(holder.desc)))

    def add_points(self, parameters: dict) -> None:
        if "new" not in parameters:
            return
        new = parameters["new"]
        pos = np.asarray(new.get("position", np.zeros((0, 3), np.float32)))
        if pos.ndim != 2 or pos.shape[1] != 3:
            raise RuntimeError("add_new_points: 'new.position' must have shape (N,3)")
        if pos.shape[0] == 0:
            return
        holder = self._points_from_dict(new, pos.shape[0], "add_new_points")
        check(self._L.pale_add_points(self._h, C.byref(holder.desc)))

    def remove_points(self, parameters: dict) -> None:
        if "indices" not in parameters:
            raise RuntimeError("remove_points: expected key 'indices' (1D int32 / int64 array)")
        idx = np.asarray(parameters["indices"])
        if idx.ndim != 1:
            raise RuntimeError("remove_points: 'indices' must be a 1D array")
        if idx.size == 0:
            return
        if idx.dtype.kind not in "iu" or idx.dtype.itemsize not in (4, 8):
            raise RuntimeError("remove_points: 'indices' must have dtype int32 or int64")
        idx64 = np.ascontiguousarray(idx, dtype=np.int64)
        check(self._L.pale_remove_points(self._h, idx64.ctypes.data_as(C.POINTER(C.c_int64)), idx64.size))

    def rebuild_bvh(self) -> None:
        check(self._L.pale_rebuild_bvh(self._h))

    def set_point_properties(self, translation3, rotation_quat4, scale3, albedo3, opacity, beta, index=-1) -> None:
        if len(translation3) != 3 or len(rotation_quat4) != 4 or len(scale3) != 3:
            raise RuntimeError("Expected translation(3), rotation_quat(4), scale(3)")
        t = np.asarray(translation3, np.float32)
        q = np.asarray(rotation_quat4, np.float32)
        s = np.asarray(scale3, np.float32)
        c = np.asarray(albedo3, np.float32)
        check(self._L.pale_set_point_properties(self._h, _p(t), _p(q), _p(s), _p(c), float(opacity), float(beta),
                                                int(index)))

    def set_point_opacity(self, opacity, index) -> None:
        check(self._L.pale_set_point_opacity(self._h, float(opacity), int(index)))

    def set_point_beta(self, beta, index) -> None:
        check(self._L.pale_set_point_beta(self._h, float(beta), int(index)))

    # -------------------------------------------------------- diagnostics --
    def trace_rays(self, origins: np.ndarray, directions: np.ndarray) -> dict:
        o = _abi.f32(origins).reshape(-1, 3)
        d = _abi.f32(directions).reshape(-1, 3)
        n = o.shape[0]
        inst = np.empty(n, np.int32)
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        pos = np.empty((n, 3), np.float32)
        alpha = np.empty(n, np.float32)
        check(self._L.pale_trace_rays(self._h, n, _p(o), _p(d), inst.ctypes.data_as(C.POINTER(C.c_int32)),
                                      prim.ctypes.data_as(C.POINTER(C.c_int32)), _p(t), _p(pos), _p(alpha)))
        return dict(instance=inst, primitive=prim, t=t, position=pos, alpha=alpha)

    def gather_irradiance(self, positions: np.ndarray, normals: np.ndarray) -> np.ndarray:
        """Irradiance estimate of the current photon map at points (n, 3)."""
        q = _abi.f32(positions).reshape(-1, 3)
        nr = _abi.f32(normals).reshape(-1, 3)
        if nr.shape != q.shape:
            raise RuntimeError(f"positions {q.shape} and normals {nr.shape} differ")
        out = np.empty_like(q)
        check(self._L.pale_gather_irradiance(self._h, q.shape[0], _p(q), _p(nr), _p(out)))
        return out

    def surfel_bvh(self) -> dict:
        cnt = C.c_uint32()
        check(self._L.pale_get_surfel_bvh(self._h, C.byref(cnt), None, None, None, None))
        n = cnt.value
        bounds = np.empty((n, 6), np.float32)
        lf = np.empty(n, np.uint32)
        c = np.empty(n, np.uint32)
        perm = np.empty(self.point_count, np.uint32)
        u = C.POINTER(C.c_uint32)
        check(self._L.pale_get_surfel_bvh(self._h, C.byref(cnt), _p(bounds), lf.ctypes.data_as(u),
                                          c.ctypes.data_as(u), perm.ctypes.data_as(u)))
        return dict(bounds=bounds, left_first=lf, count=c, permutation=perm)

    def photons(self) -> dict:
        cnt = C.c_uint32()
        att = C.c_uint64()
        check(self._L.pale_get_photons(self._h, C.byref(cnt), C.byref(att), None, None, None))
        n = cnt.value
        pos = np.empty((n, 3), np.float32)
        pw = np.empty((n, 3), np.float32)
        dr = np.empty((n, 3), np.float32)
        check(self._L.pale_get_photons(self._h, C.byref(cnt), C.byref(att), _p(pos), _p(pw), _p(dr)))
        return dict(count=n, attempted=att.value, position=pos, power=pw, direction=dr)

    def set_count_work(self, on: bool) -> None:
        check(self._L.pale_set_count_work(self._h, 1 if on else 0))

    def stats(self) -> dict:
        s = _abi.Stats()
        check(self._L.pale_get_stats(self._h, C.byref(s)))
        return s.as_dict()
