"""ctypes mirror of include/pale_abi.h and helpers to marshal scenes.

The structs here must match include/pale_abi.h field for field; the test
``tests/test_abi_layout.py`` checks their sizes against the C compiler's.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

PALE_NAME_LEN = 16

PALE_OK = 0
PALE_ERR_INVALID_ARG = -1
PALE_ERR_SIZE = -2
PALE_ERR_RANGE = -3
PALE_ERR_DEVICE = -4
PALE_ERR_INTERNAL = -5

_fp = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)


class MeshDesc(C.Structure):
    _fields_ = [
        ("positions", _fp),
        ("vertex_count", C.c_uint32),
        ("indices", _u32p),
        ("triangle_count", C.c_uint32),
        ("object_to_world", C.c_float * 16),
        ("world_to_object", C.c_float * 16),
        ("base_color", C.c_float * 3),
        ("power", C.c_float),
        ("name", C.c_char * PALE_NAME_LEN),
    ]


class CameraDesc(C.Structure):
    _fields_ = [
        ("name", C.c_char * PALE_NAME_LEN),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("fovy_deg", C.c_float),
        ("view", C.c_float * 16),
        ("inv_view", C.c_float * 16),
        ("position", C.c_float * 3),
        ("forward", C.c_float * 3),
        ("use_for_adjoint", C.c_int32),
    ]


class BatchTuning(C.Structure):
    _fields_ = [
        ("slots_per_batch", C.c_uint32),
        ("entries_per_slot", C.c_uint32),
        ("graph_rounds", C.c_uint32),
        ("grid_fraction", C.c_float),
        ("walk_lds_entries", C.c_uint32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PointsDesc(C.Structure):
    _fields_ = [
        ("count", C.c_uint32),
        ("position", _fp),
        ("tangent_u", _fp),
        ("tangent_v", _fp),
        ("scale", _fp),
        ("albedo", _fp),
        ("opacity", _fp),
        ("beta", _fp),
        ("shape", _fp),
    ]


class SceneDesc(C.Structure):
    _fields_ = [
        ("meshes", C.POINTER(MeshDesc)),
        ("mesh_count", C.c_uint32),
        ("cameras", C.POINTER(CameraDesc)),
        ("camera_count", C.c_uint32),
        ("points", PointsDesc),
    ]


class Settings(C.Structure):
    _fields_ = [
        ("photons", C.c_uint32),
        ("bounces", C.c_uint32),
        ("forward_passes", C.c_uint32),
        ("gather_passes", C.c_uint32),
        ("adjoint_bounces", C.c_uint32),
        ("adjoint_passes", C.c_uint32),
        ("seed", C.c_uint64),
        ("gather_radius", C.c_float),
        ("rr_start", C.c_uint32),
        ("q_null", C.c_float),
        ("q_reflect", C.c_float),
        ("q_transmit", C.c_float),
        ("spp_decorrelate", C.c_int32),
        ("debug_images", C.c_int32),
        ("gather_entry_budget", C.c_uint32),
        ("photon_map_capacity", C.c_uint32),
        ("full_gradients", C.c_int32),
        ("batch_passes", C.c_int32),
        ("adjoint_decorrelate", C.c_int32),
        ("adjoint_seed", C.c_uint32),
        ("bvh_build", C.c_uint32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("photon_queries", C.c_uint64),
        ("gather_queries", C.c_uint64),
        ("adjoint_queries", C.c_uint64),
        ("photons_deposited", C.c_uint64),
        ("gradient_events", C.c_uint64),
        ("grid_cells", C.c_uint64),
        ("forward_ms", C.c_double),
        ("backward_ms", C.c_double),
        ("photon_trace_ms", C.c_double),
        ("gather_ms", C.c_double),
        ("adjoint_trace_ms", C.c_double),
        ("adjoint_transport_ms", C.c_double),
        ("grid_ms", C.c_double),
        ("photon_trace_launches", C.c_uint32),
        ("gather_launches", C.c_uint32),
        ("adjoint_trace_launches", C.c_uint32),
        ("adjoint_transport_launches", C.c_uint32),
        ("camera_trace_ms", C.c_double),
        ("camera_sort_ms", C.c_double),
        ("camera_trace_launches", C.c_uint32),
        ("camera_rounds", C.c_uint32),
        ("work", (C.c_uint64 * 7) * 3),
        ("bvh_build_ms", C.c_double),
        ("bvh_on_device", C.c_uint32),
        ("bvh_pad", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        d = {name: getattr(self, name) for name, _ in self._fields_ if name not in ("work", "bvh_pad")}
        for ci, cname in enumerate(WORK_CLASSES):
            d[f"work_{cname}"] = {u: int(self.work[ci][k]) for k, u in enumerate(WORK_UNITS)}
        return d


WORK_CLASSES = ("photon", "gather", "adjoint")
WORK_UNITS = ("pops", "aabb", "prims", "gathers", "cells", "cells_nonempty", "photons")


# Reference defaults: PathTracerSettings (GPUDataStructures.h:400-422),
# gather radius hard-coded at PathTracer.cpp:187.
DEFAULT_SETTINGS = dict(
    photons=1_000_000,
    bounces=6,
    forward_passes=6,
    gather_passes=6,
    adjoint_bounces=6,
    adjoint_passes=6,
    seed=42,
    gather_radius=0.01,
    rr_start=6,
    q_null=0.5,
    q_reflect=0.5,
    q_transmit=0.0,
    spp_decorrelate=0,
    debug_images=0,
    gather_entry_budget=0,
    photon_map_capacity=0,
    full_gradients=0,
    batch_passes=1,
    adjoint_decorrelate=0,
    adjoint_seed=42,
    bvh_build=0,
)


def make_settings(d: Optional[dict] = None) -> Settings:
    """Settings struct from a reference-style settings dict.

    Unknown keys are ignored, as in pale_bindings.cpp:123-139."""
    vals = dict(DEFAULT_SETTINGS)
    for k, v in (d or {}).items():
        if k in vals:
            vals[k] = v
    s = Settings()
    for k, v in vals.items():
        if k in ("gather_radius", "q_null", "q_reflect", "q_transmit"):
            setattr(s, k, float(v))
        elif k in ("spp_decorrelate", "debug_images", "full_gradients", "batch_passes", "adjoint_decorrelate"):
            setattr(s, k, int(bool(v)))
        else:
            setattr(s, k, int(v))
    return s


def f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def fptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_fp)


@dataclass
class Mesh:
    positions: np.ndarray          # (V,3) object space
    indices: np.ndarray            # (T,3)
    object_to_world: np.ndarray    # (4,4)
    world_to_object: np.ndarray    # (4,4)
    base_color: Sequence[float]
    power: float
    name: str = "mesh"


@dataclass
class Camera:
    name: str
    width: int
    height: int
    fovy_deg: float
    view: np.ndarray      # (4,4) world -> camera
    inv_view: np.ndarray  # (4,4) camera -> world
    position: Sequence[float]
    forward: Sequence[float]
    use_for_adjoint: bool = True


POINT_KEYS = ("position", "tangent_u", "tangent_v", "scale", "albedo", "opacity", "beta", "shape")
POINT_SHAPES = {"position": 3, "tangent_u": 3, "tangent_v": 3, "scale": 2, "albedo": 3,
                "opacity": 0, "beta": 0, "shape": 0}


class PointsHolder:
    """Keeps numpy arrays alive while a PointsDesc points at them."""

    def __init__(self, points: dict, require_shape: bool = False):
        n = int(np.asarray(points["position"]).shape[0])
        self.arrays = {}
        for k in POINT_KEYS:
            if k not in points or points[k] is None:
                if k == "shape" and not require_shape:
                    self.arrays[k] = None
                    continue
                raise KeyError(k)
            a = f32(points[k])
            w = POINT_SHAPES[k]
            want = (n, w) if w else (n,)
            if a.shape != want:
                raise ValueError(f"'{k}' must have shape {want}, got {a.shape}")
            self.arrays[k] = a
        self.desc = PointsDesc()
        self.desc.count = n
        for k in POINT_KEYS:
            setattr(self.desc, k, fptr(self.arrays[k]))


def fill_camera_desc(d: CameraDesc, c: Camera) -> CameraDesc:
    d.name = c.name.encode()[: PALE_NAME_LEN - 1]
    d.width = int(c.width)
    d.height = int(c.height)
    d.fovy_deg = float(np.float32(c.fovy_deg))
    d.view[:] = [float(x) for x in f32(c.view).reshape(16)]
    d.inv_view[:] = [float(x) for x in f32(c.inv_view).reshape(16)]
    d.position[:] = [float(x) for x in f32(c.position).reshape(3)]
    d.forward[:] = [float(x) for x in f32(c.forward).reshape(3)]
    d.use_for_adjoint = 1 if c.use_for_adjoint else 0
    return d


class SceneHolder:
    """Builds a SceneDesc and keeps every buffer it references alive."""

    def __init__(self, meshes: List[Mesh], cameras: List[Camera], points: dict):
        self.meshes = meshes
        self.cameras = cameras
        self._keep = []
        mesh_arr = (MeshDesc * max(1, len(meshes)))()
        for i, m in enumerate(meshes):
            pos = f32(m.positions).reshape(-1, 3)
            idx = np.ascontiguousarray(np.asarray(m.indices, dtype=np.uint32).reshape(-1, 3))
            self._keep += [pos, idx]
            d = mesh_arr[i]
            d.positions = pos.ctypes.data_as(_fp)
            d.vertex_count = pos.shape[0]
            d.indices = idx.ctypes.data_as(_u32p)
            d.triangle_count = idx.shape[0]
            d.object_to_world[:] = [float(x) for x in f32(m.object_to_world).reshape(16)]
            d.world_to_object[:] = [float(x) for x in f32(m.world_to_object).reshape(16)]
            d.base_color[:] = [float(x) for x in f32(m.base_color).reshape(3)]
            d.power = float(np.float32(m.power))
            d.name = m.name.encode()[: PALE_NAME_LEN - 1]
        cam_arr = (CameraDesc * max(1, len(cameras)))()
        for i, c in enumerate(cameras):
            fill_camera_desc(cam_arr[i], c)
        self.points = PointsHolder(points)
        self._mesh_arr = mesh_arr
        self._cam_arr = cam_arr
        self.desc = SceneDesc()
        self.desc.meshes = C.cast(mesh_arr, C.POINTER(MeshDesc))
        self.desc.mesh_count = len(meshes)
        self.desc.cameras = C.cast(cam_arr, C.POINTER(CameraDesc))
        self.desc.camera_count = len(cameras)
        self.desc.points = self.points.desc
