"""Scene and camera ingestion: 3D Gaussian Splatting PLY files and
``cameras.json`` camera lists, in the conventions the rasterizer reads.

The reference receives its tensors from an external framework and has no
on-disk format. The de-facto interchange format of Gaussian splatting is the
PLY that 3DGS training writes: one ``vertex`` element per Gaussian with float
properties ``x y z`` (mean), ``f_dc_0..2`` (degree-0 SH per channel),
``f_rest_0..(3*(B-1)-1)`` (higher SH, channel-major: all coefficients of red,
then green, then blue), ``opacity`` (logit), ``scale_0..2`` (log standard
deviations) and ``rot_0..3`` (quaternion w, x, y, z, not normalised); other
properties (normals) are ignored. ``read_ply`` returns the rasterizer's
activated layout: linear scales, sigmoid opacity, unit quaternions,
``sh_0 (N,1,3)`` and ``sh_rest (N,B-1,3)``. ``write_ply`` is its inverse.

``cameras.json`` (as written by 3DGS): a list of objects with ``width``,
``height``, ``fx``, ``fy``, ``position`` (camera centre in world space) and
``rotation`` (3x3 camera-to-world rotation; its columns are the camera's x
right, y down, z forward axes), the principal point at the image centre.
"""
from __future__ import annotations

import json
import math

import numpy as np

from .scene import Camera, camera_from_intrinsics

_PLY_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2", "int16": "i2",
    "ushort": "u2", "uint16": "u2", "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
    "float": "f4", "float32": "f4", "double": "f8", "float64": "f8",
}


def _parse_header(f) -> tuple:
    if f.readline().strip() != b"ply":
        raise ValueError("not a PLY file")
    fmt, n, props, in_vertex, seen_vertex = None, None, [], False, False
    while True:
        line = f.readline()
        if not line:
            raise ValueError("PLY header ends without end_header")
        tok = line.decode("ascii", errors="replace").split()
        if not tok or tok[0] == "comment" or tok[0] == "obj_info":
            continue
        if tok[0] == "end_header":
            break
        if tok[0] == "format":
            fmt = tok[1]
        elif tok[0] == "element":
            if seen_vertex and in_vertex:
                in_vertex = False
            if tok[1] == "vertex":
                if seen_vertex:
                    raise ValueError("two vertex elements")
                n, in_vertex, seen_vertex = int(tok[2]), True, True
            elif not seen_vertex:
                raise ValueError("elements before the vertex element are not supported")
            else:
                in_vertex = False
        elif tok[0] == "property" and in_vertex:
            if tok[1] == "list":
                raise ValueError("list properties in the vertex element are not supported")
            if tok[1] not in _PLY_TYPES:
                raise ValueError(f"unknown PLY type {tok[1]}")
            props.append((tok[2], _PLY_TYPES[tok[1]]))
    if fmt not in ("binary_little_endian", "binary_big_endian", "ascii"):
        raise ValueError(f"unsupported PLY format {fmt}")
    if n is None:
        raise ValueError("no vertex element")
    return fmt, n, props


def read_ply(path: str) -> dict:
    """Gaussians of a 3DGS PLY file as the rasterizer's scene arrays (float32)."""
    with open(path, "rb") as f:
        fmt, n, props = _parse_header(f)
        names = [p for p, _ in props]
        if fmt == "ascii":
            rows = np.loadtxt(f, dtype=np.float64, max_rows=n, ndmin=2)
            if rows.shape != (n, len(props)):
                raise ValueError(f"expected {n} rows of {len(props)} values")
            col = {p: rows[:, k] for k, p in enumerate(names)}
        else:
            end = "<" if fmt == "binary_little_endian" else ">"
            dt = np.dtype([(p, end + t) for p, t in props])
            data = np.fromfile(f, dtype=dt, count=n)
            if data.shape[0] != n:
                raise ValueError(f"file holds {data.shape[0]} of {n} vertices")
            col = {p: data[p] for p in names}
    need = ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1", "scale_2",
            "rot_0", "rot_1", "rot_2", "rot_3"]
    missing = [p for p in need if p not in col]
    if missing:
        raise ValueError(f"PLY lacks Gaussian properties {missing}")
    rest = sorted((p for p in names if p.startswith("f_rest_")), key=lambda p: int(p[7:]))
    if [int(p[7:]) for p in rest] != list(range(len(rest))) or len(rest) % 3:
        raise ValueError("f_rest_* properties must be f_rest_0 .. f_rest_{3k-1}")
    k = len(rest) // 3
    if (k + 1) not in (1, 4, 9, 16):
        raise ValueError(f"{k + 1} SH bases per channel is not a complete degree 0..3")
    f64 = lambda p: np.asarray(col[p], dtype=np.float64)
    pos = np.stack([f64("x"), f64("y"), f64("z")], axis=1)
    scales = np.exp(np.stack([f64("scale_0"), f64("scale_1"), f64("scale_2")], axis=1))
    q = np.stack([f64("rot_0"), f64("rot_1"), f64("rot_2"), f64("rot_3")], axis=1)
    qn = np.linalg.norm(q, axis=1, keepdims=True)
    q = np.where(qn > 0, q / np.where(qn > 0, qn, 1.0), np.array([1.0, 0.0, 0.0, 0.0]))
    opa = 1.0 / (1.0 + np.exp(-f64("opacity")))
    sh0 = np.stack([f64("f_dc_0"), f64("f_dc_1"), f64("f_dc_2")], axis=1).reshape(n, 1, 3)
    if k:
        shr = np.stack([f64(p) for p in rest], axis=1).reshape(n, 3, k).transpose(0, 2, 1)
    else:
        shr = np.zeros((n, 0, 3))
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return {"positions": f32(pos), "scales": f32(scales), "rotations": f32(q), "opacities": f32(opa),
            "sh_0": f32(sh0), "sh_rest": f32(shr)}


def write_ply(path: str, scene: dict) -> None:
    """Writes scene arrays (activated, as read_ply returns them) as a binary
    little-endian 3DGS PLY."""
    pos = np.asarray(scene["positions"], dtype=np.float64)
    n = pos.shape[0]
    shr = np.asarray(scene["sh_rest"], dtype=np.float64).reshape(n, -1, 3)
    k = shr.shape[1]
    opa = np.clip(np.asarray(scene["opacities"], dtype=np.float64).reshape(n), 1e-12, 1.0 - 1e-7)
    cols = {"x": pos[:, 0], "y": pos[:, 1], "z": pos[:, 2],
            "nx": np.zeros(n), "ny": np.zeros(n), "nz": np.zeros(n)}
    sh0 = np.asarray(scene["sh_0"], dtype=np.float64).reshape(n, 3)
    for c in range(3):
        cols[f"f_dc_{c}"] = sh0[:, c]
    rest = shr.transpose(0, 2, 1).reshape(n, 3 * k)  # channel-major
    for j in range(3 * k):
        cols[f"f_rest_{j}"] = rest[:, j]
    cols["opacity"] = np.log(opa / (1.0 - opa))
    sc = np.log(np.asarray(scene["scales"], dtype=np.float64))
    for c in range(3):
        cols[f"scale_{c}"] = sc[:, c]
    q = np.asarray(scene["rotations"], dtype=np.float64)
    for c in range(4):
        cols[f"rot_{c}"] = q[:, c]
    names = list(cols)
    dt = np.dtype([(p, "<f4") for p in names])
    data = np.empty(n, dtype=dt)
    for p in names:
        data[p] = cols[p]
    header = ["ply", "format binary_little_endian 1.0", f"element vertex {n}"]
    header += [f"property float {p}" for p in names] + ["end_header"]
    with open(path, "wb") as f:
        f.write(("\n".join(header) + "\n").encode("ascii"))
        data.tofile(f)


def camera_from_json(entry: dict, near: float = 0.2, far: float = 1000.0) -> Camera:
    """One cameras.json entry as a rasterizer Camera."""
    w, h = int(entry["width"]), int(entry["height"])
    c2w = np.asarray(entry["rotation"], dtype=np.float64).reshape(3, 3)
    eye = np.asarray(entry["position"], dtype=np.float64).reshape(3)
    fx, fy = float(entry["fx"]), float(entry["fy"])
    cx = float(entry.get("cx", w / 2.0))
    cy = float(entry.get("cy", h / 2.0))
    return camera_from_intrinsics(w, h, fx, fy, cx, cy, eye, c2w.T, near, far)


def camera_to_json(cam: Camera, ident: int = 0, name: str = "") -> dict:
    """The inverse of camera_from_json."""
    R = np.asarray(cam.M, dtype=np.float64)[:3, :3]
    return {"id": ident, "img_name": name, "width": cam.width, "height": cam.height,
            "position": [float(v) for v in cam.position], "rotation": R.T.tolist(),
            "fx": float(cam.fx), "fy": float(cam.fy)}


def read_cameras_json(path: str, near: float = 0.2, far: float = 1000.0) -> list:
    with open(path) as f:
        entries = json.load(f)
    if isinstance(entries, dict):
        entries = [entries]
    return [camera_from_json(e, near, far) for e in entries]


def fov_x_deg(cam: Camera) -> float:
    return math.degrees(2.0 * math.atan(0.5 * cam.width / cam.fx))
