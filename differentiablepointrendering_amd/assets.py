"""Scene ingestion for ``pale.Renderer(assetRootDir, sceneXml, pointCloudFile)``.

Reads the reference's inputs on the host side of the C ABI:

* the Mitsuba-XML subset of SceneSerializer.cpp:129-377 -- perspective
  sensors (fov / fov_axis, lookat, film width/height, ``adjoint_source``
  flag), shapes of type obj / ply with a ``to_world`` of translate / rotate /
  scale (the last of each kind wins, composed T * R * S as
  TransformComponent::getTransform, Components.ixx:68-73), diffuse bsdf
  reflectance by reference, and area emitters (colour + power,
  MaterialBakerMitsuba.ixx:42-66);
* meshes in Wavefront OBJ (polygons fan-triangulated) or PLY;
* the surfel PLY of PLYPointLoader.ixx:158-383: the beta-surfel schema
  (x y z, tu_*, tv_*, su sv, albedo_*, opacity, beta, shape; frames
  canonicalised face-up) or the 2DGS schema (f_dc_*, opacity logit,
  scale_* log, rot_*), ascii or binary.

Matrices are computed in float64 and rounded to float32 (the reference uses
glm in float32; that rounding is not reproduced).
"""
from __future__ import annotations

import math
import os
import re
import xml.etree.ElementTree as ET
from typing import Dict, List, Tuple

import numpy as np

from ._abi import Camera, Mesh

# --------------------------------------------------------------------- PLY --
_PLY_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2", "int16": "i2", "ushort": "u2",
    "uint16": "u2", "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4", "float": "f4", "float32": "f4",
    "double": "f8", "float64": "f8",
}


def read_ply(path: str) -> Dict[str, Dict[str, np.ndarray]]:
    """Elements of a PLY file: {element: {property: array}}; list
    properties (faces) come back as a list of index arrays."""
    with open(path, "rb") as f:
        data = f.read()
    end = data.find(b"end_header")
    if not data.startswith(b"ply") or end < 0:
        raise RuntimeError(f"{path}: not a PLY file")
    nl = data.find(b"\n", end)
    header = data[:end].decode("ascii", "replace").splitlines()
    body = data[nl + 1:]
    fmt = None
    elements: List[Tuple[str, int, list]] = []
    for line in header:
        tok = line.split("#")[0].split()
        if not tok:
            continue
        if tok[0] == "format":
            fmt = tok[1]
        elif tok[0] == "element":
            elements.append((tok[1], int(tok[2]), []))
        elif tok[0] == "property":
            if tok[1] == "list":
                elements[-1][2].append((tok[4], ("list", _PLY_TYPES[tok[2]], _PLY_TYPES[tok[3]])))
            else:
                elements[-1][2].append((tok[2], _PLY_TYPES[tok[1]]))
    out: Dict[str, Dict[str, np.ndarray]] = {}
    if fmt == "ascii":
        words = body.decode("ascii", "replace").split()
        pos = 0
        for name, count, props in elements:
            cols = {p[0]: [] for p in props}
            for _ in range(count):
                for pname, ptype in props:
                    if isinstance(ptype, tuple):
                        k = int(words[pos]); pos += 1
                        cols[pname].append(np.array(words[pos:pos + k], dtype=ptype[2])); pos += k
                    else:
                        cols[pname].append(words[pos]); pos += 1
            out[name] = {p[0]: (cols[p[0]] if isinstance(p[1], tuple) else np.array(cols[p[0]], dtype=p[1]))
                         for p in props}
        return out
    if fmt not in ("binary_little_endian", "binary_big_endian"):
        raise RuntimeError(f"{path}: unsupported PLY format {fmt}")
    endian = "<" if fmt == "binary_little_endian" else ">"
    pos = 0
    for name, count, props in elements:
        if all(not isinstance(p[1], tuple) for p in props):
            dt = np.dtype([(p[0], endian + p[1]) for p in props])
            arr = np.frombuffer(body, dtype=dt, count=count, offset=pos)
            pos += dt.itemsize * count
            out[name] = {p[0]: np.array(arr[p[0]]) for p in props}
        else:
            cols = {p[0]: [] for p in props}
            for _ in range(count):
                for pname, ptype in props:
                    if isinstance(ptype, tuple):
                        ct, it = np.dtype(endian + ptype[1]), np.dtype(endian + ptype[2])
                        k = int(np.frombuffer(body, ct, 1, pos)[0]); pos += ct.itemsize
                        cols[pname].append(np.frombuffer(body, it, k, pos).copy()); pos += it.itemsize * k
                    else:
                        t = np.dtype(endian + ptype)
                        cols[pname].append(np.frombuffer(body, t, 1, pos)[0]); pos += t.itemsize
            out[name] = {p[0]: (cols[p[0]] if isinstance(p[1], tuple) else np.array(cols[p[0]]))
                         for p in props}
    return out


def _normalize(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def canonicalize_face_up(tu: np.ndarray, tv: np.ndarray):
    """canonicalizeFrameFaceUp (PLYPointLoader.ixx:75-101) in float32:
    Gram-Schmidt, then flip the second tangent so that the normal has z >= 0."""
    tu = _normalize(tu.astype(np.float32))
    tv = _normalize(tv.astype(np.float32))
    ortho = tv - np.sum(tv * tu, -1, keepdims=True) * tu
    bad = np.sum(ortho * ortho, -1) < 1e-6
    if bad.any():
        alt = np.where((np.abs(tu[bad, 1]) < 0.9)[:, None], np.array([0, 1, 0], np.float32),
                       np.array([1, 0, 0], np.float32))
        ortho[bad] = alt - np.sum(alt * tu[bad], -1, keepdims=True) * tu[bad]
    tv = _normalize(ortho)
    n = np.cross(tu, tv)
    flip = n[:, 2] < 0
    tv[flip] = -tv[flip]
    return tu.astype(np.float32), tv.astype(np.float32)


def load_points(path: str) -> Dict[str, np.ndarray]:
    """Surfel PLY -> point dict (asset order)."""
    v = read_ply(path).get("vertex")
    if v is None or len(v.get("x", [])) == 0:
        raise RuntimeError(f"PLYPointLoader: zero vertices in {path}")
    f = lambda *ks: np.stack([np.asarray(v[k], np.float32) for k in ks], axis=1)
    n = len(v["x"])
    custom = all(k in v for k in ("x", "y", "z", "tu_x", "tu_y", "tu_z", "tv_x", "tv_y", "tv_z", "su", "sv",
                                  "albedo_r", "albedo_g", "albedo_b", "opacity"))
    gs = all(k in v for k in ("x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1", "rot_0",
                              "rot_1", "rot_2", "rot_3"))
    if custom:
        tu, tv = canonicalize_face_up(f("tu_x", "tu_y", "tu_z"), f("tv_x", "tv_y", "tv_z"))
        return dict(position=f("x", "y", "z"), tangent_u=tu, tangent_v=tv, scale=f("su", "sv"),
                    albedo=np.clip(f("albedo_r", "albedo_g", "albedo_b"), 0, 1),
                    opacity=np.asarray(v["opacity"], np.float32),
                    beta=np.asarray(v.get("beta", np.zeros(n)), np.float32),
                    shape=np.asarray(v.get("shape", np.zeros(n)), np.float32))
    if gs:
        q = f("rot_0", "rot_1", "rot_2", "rot_3").astype(np.float64)  # (w, x, y, z)
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        w, x, y, z = q.T
        # The reference reads row 0 and row 1 of transpose(mat3_cast(q)) in
        # glm's column-major indexing: the first two columns of R(q).
        tu = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)], 1)
        tv = np.stack([2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)], 1)
        tu = _normalize(tu)
        tv = _normalize(tv - np.sum(tv * tu, 1, keepdims=True) * tu)
        scale = np.exp(np.minimum(f("scale_0", "scale_1").astype(np.float64), 80.0)) * 2.0
        op = 1.0 / (1.0 + np.exp(-np.asarray(v["opacity"], np.float64)))
        return dict(position=f("x", "y", "z"), tangent_u=tu.astype(np.float32), tangent_v=tv.astype(np.float32),
                    scale=scale.astype(np.float32), albedo=np.full((n, 3), 0.8, np.float32),
                    opacity=op.astype(np.float32), beta=np.zeros(n, np.float32), shape=np.zeros(n, np.float32))
    raise RuntimeError(f"PLYPointLoader: unsupported vertex schema in {path}")


def save_points(path: str, points: Dict[str, np.ndarray]) -> None:
    """Write a point dict as an ascii beta-surfel PLY (the custom schema)."""
    keys = [("x", "position", 0), ("y", "position", 1), ("z", "position", 2), ("tu_x", "tangent_u", 0),
            ("tu_y", "tangent_u", 1), ("tu_z", "tangent_u", 2), ("tv_x", "tangent_v", 0), ("tv_y", "tangent_v", 1),
            ("tv_z", "tangent_v", 2), ("su", "scale", 0), ("sv", "scale", 1), ("albedo_r", "albedo", 0),
            ("albedo_g", "albedo", 1), ("albedo_b", "albedo", 2), ("opacity", "opacity", None),
            ("beta", "beta", None), ("shape", "shape", None)]
    n = len(points["position"])
    cols = []
    for _, k, i in keys:
        a = np.asarray(points.get(k, np.zeros(n)), np.float32)
        cols.append(a[:, i] if i is not None else a)
    table = np.stack(cols, 1)
    with open(path, "w") as f:
        f.write("ply\nformat ascii 1.0\nelement vertex %d\n" % n)
        for name, _, _ in keys:
            f.write(f"property float {name}\n")
        f.write("end_header\n")
        np.savetxt(f, table, fmt="%.9g")


# ------------------------------------------------------------------ meshes --
def read_mesh(path: str) -> Tuple[np.ndarray, np.ndarray]:
    """(positions (V,3), triangles (T,3)) of an OBJ or PLY mesh."""
    if path.lower().endswith(".obj"):
        verts, tris = [], []
        with open(path) as f:
            for line in f:
                tok = line.split()
                if not tok:
                    continue
                if tok[0] == "v":
                    verts.append([float(t) for t in tok[1:4]])
                elif tok[0] == "f":
                    idx = []
                    for t in tok[1:]:
                        i = int(t.split("/")[0])
                        idx.append(i - 1 if i > 0 else len(verts) + i)
                    for k in range(1, len(idx) - 1):
                        tris.append([idx[0], idx[k], idx[k + 1]])
        return np.asarray(verts, np.float32).reshape(-1, 3), np.asarray(tris, np.uint32).reshape(-1, 3)
    e = read_ply(path)
    v = e["vertex"]
    pos = np.stack([np.asarray(v[k], np.float32) for k in ("x", "y", "z")], 1)
    faces = e.get("face", {})
    lst = faces.get("vertex_indices", faces.get("vertex_index", []))
    tris = []
    for fi in lst:
        for k in range(1, len(fi) - 1):
            tris.append([fi[0], fi[k], fi[k + 1]])
    return pos, np.asarray(tris, np.uint32).reshape(-1, 3)


# ------------------------------------------------------------------- XML --
def _floats(s: str, n: int = 3) -> List[float]:
    vals = [float(x) for x in re.split(r"[,\s]+", s.strip()) if x]
    return (vals + [0.0] * n)[:n]


def _child_by_name(node, tag: str, name: str):
    for c in node.findall(tag):
        if c.get("name") == name:
            return c
    return None


def rotation_matrix(axis, angle_deg: float) -> np.ndarray:
    a = np.asarray(axis, np.float64)
    if np.linalg.norm(a) > 0:
        a = a / np.linalg.norm(a)
    t = math.radians(angle_deg)
    c, s = math.cos(t), math.sin(t)
    x, y, z = a
    return np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                     [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                     [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]])


def shape_transform(node) -> np.ndarray:
    """to_world of a shape: the last translate, rotate and scale, composed
    translate * rotate * scale (SceneSerializer.cpp:282-303)."""
    T, R, S = np.zeros(3), np.eye(3), np.ones(3)
    tw = _child_by_name(node, "transform", "to_world")
    if tw is not None:
        for c in tw:
            if c.tag == "translate":
                T = np.array([float(c.get(k, 0)) for k in "xyz"])
            elif c.tag == "scale":
                S = np.array([float(c.get(k, 1)) for k in "xyz"])
            elif c.tag == "rotate":
                R = rotation_matrix([float(c.get(k, 0)) for k in "xyz"], float(c.get("angle", 0)))
    M = np.eye(4)
    M[:3, :3] = R @ np.diag(S)
    M[:3, 3] = T
    return M


def fov_y_degrees(fov: float, axis: str, w: int, h: int) -> float:
    """computeFovYDegrees (SceneSerializer.cpp:71-86)."""
    aspect = w / h
    if axis == "y":
        return fov
    if axis == "x" or w <= h:
        return math.degrees(2.0 * math.atan(math.tan(math.radians(fov) * 0.5) / aspect))
    return fov


def _bool(s: str, default=True) -> bool:
    s = (s or "").strip().lower()
    if s in ("true", "1", "yes", "on"):
        return True
    if s in ("false", "0", "no", "off"):
        return False
    try:
        return float(s) != 0.0
    except ValueError:
        return default


def load_xml(xml_path: str, asset_root: str):
    """Cameras and meshes of a scene XML (paths relative to the asset root)."""
    root = ET.parse(xml_path).getroot()
    if root.tag != "scene":
        raise RuntimeError(f"{xml_path}: missing <scene> root")
    cams: List[Camera] = []
    for s in root.findall("sensor"):
        name = s.get("id", "")
        film = s.find("film")
        w = int(_child_by_name(film, "integer", "width").get("value"))
        h = int(_child_by_name(film, "integer", "height").get("value"))
        fov_n = _child_by_name(s, "float", "fov")
        fov = float(fov_n.get("value")) if fov_n is not None else 45.0
        ax_n = _child_by_name(s, "string", "fov_axis")
        axis = ax_n.get("value") if ax_n is not None else "smaller"
        la = s.find("transform").find("lookat")
        eye, tgt, up = (_floats(la.get(k, "0,0,0")) for k in ("origin", "target", "up"))
        adj = True
        for tag in ("boolean", "float"):
            n = _child_by_name(s, tag, "adjoint_source")
            if n is not None:
                adj = _bool(n.get("value"), True)
        from .scene import look_at_camera

        cams.append(look_at_camera(name, eye, tgt, up, w, h, fov_y_degrees(fov, axis, w, h), adj))
    bsdfs = {}
    for b in root.findall("bsdf"):
        rgb = _child_by_name(b, "rgb", "reflectance")
        bsdfs[b.get("id", "")] = _floats(rgb.get("value")) if rgb is not None else [0.5, 0.5, 0.5]
    meshes: List[Mesh] = []
    for sh in root.findall("shape"):
        typ = sh.get("type", "")
        if typ not in ("obj", "ply", "plymesh", "serialized"):
            continue
        fn = _child_by_name(sh, "string", "filename")
        if fn is None:
            continue
        p = fn.get("value")
        p = p if os.path.isabs(p) else os.path.join(asset_root, p)
        pos, tris = read_mesh(p)
        color, power = [0.5, 0.5, 0.5], 0.0
        em = sh.find("emitter")
        ref = sh.find("ref")
        inline = sh.find("bsdf")
        if em is not None and em.get("type", "area") == "area":
            c = _child_by_name(em, "rgb", "color")
            color = _floats(c.get("value")) if c is not None else [0.5, 0.5, 0.5]
            pw = _child_by_name(em, "float", "power")
            power = float(pw.get("value")) if pw is not None else 0.0
        elif ref is not None and ref.get("id") in bsdfs:
            color = bsdfs[ref.get("id")]
        elif inline is not None:
            rgb = _child_by_name(inline, "rgb", "reflectance")
            color = _floats(rgb.get("value")) if rgb is not None else color
        M = shape_transform(sh)
        meshes.append(Mesh(pos, tris, M.astype(np.float32), np.linalg.inv(M).astype(np.float32), color, power,
                           sh.get("id", "mesh")[:15]))
    return meshes, cams


def load_scene(asset_root: str, scene_xml: str, point_cloud: str):
    """Renderer(assetRootDir, sceneXml, pointCloudFile) inputs
    (pale_bindings.cpp:61-101): the XML relative to the asset root, the
    point cloud under <root>/PointClouds/ unless absolute."""
    xml_path = scene_xml if os.path.isabs(scene_xml) else os.path.join(asset_root, scene_xml)
    meshes, cams = load_xml(xml_path, asset_root)
    pc = point_cloud or "initial.ply"
    ply = pc if os.path.isabs(pc) else os.path.join(asset_root, "PointClouds", pc)
    return meshes, cams, load_points(ply)
