"""Particles filling a mesh: a jittered lattice kept where the mesh's winding number says inside,
the inside test on the GPU (shapes.py:1-51).

Same names and arguments as the reference module.  Differences, each stated where it applies:
the mesh is read by `read_obj` (a Wavefront OBJ reader of this package) instead of trimesh; the
inside test is `hfm_mesh_inside` (exact winding number, include/hfm.h) instead of Warp's
fast-winding-number query; no mesh ships with this package, so `BUNNY` names a file the user
provides (the reference's assets/bunny.obj).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _abi

ASSETS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "assets")
BUNNY = os.environ.get("HFM_BUNNY", os.path.join(ASSETS, "bunny.obj"))


class Mesh:
    """The two attributes of a trimesh.Trimesh that the seeding reads."""

    def __init__(self, vertices, faces):
        self.vertices = np.asarray(vertices, dtype=float).reshape(-1, 3)
        self.faces = np.asarray(faces, dtype=np.int64).reshape(-1, 3)

    def triangles(self) -> np.ndarray:
        """(F, 9) float32: the three vertices of every face."""
        return np.ascontiguousarray(self.vertices[self.faces].reshape(-1, 9), dtype=np.float32)


def read_obj(path: str) -> Mesh:
    """Vertices ('v x y z') and faces ('f a b c ...', 1-based or negative indices, optional
    /vt/vn parts, polygons split into fans) of a Wavefront OBJ file.  Faces with a repeated
    vertex or zero area are dropped: they enclose nothing and their solid angle is undefined."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path}: no mesh is bundled with this package; pass the path of an OBJ file "
                                "(the reference's assets/bunny.obj, or set HFM_BUNNY)")
    verts, faces = [], []
    with open(path, "r", encoding="utf-8", errors="replace") as fh:
        for line in fh:
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "v":
                verts.append([float(x) for x in parts[1:4]])
            elif parts[0] == "f":
                idx = []
                for tok in parts[1:]:
                    k = int(tok.split("/")[0])
                    idx.append(k - 1 if k > 0 else len(verts) + k)
                for q in range(1, len(idx) - 1):
                    faces.append([idx[0], idx[q], idx[q + 1]])
    v = np.asarray(verts, dtype=float).reshape(-1, 3)
    f = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    if len(f) and (f.min() < 0 or f.max() >= len(v)):
        raise ValueError(f"{path}: face index out of range")
    if len(f):
        distinct = (f[:, 0] != f[:, 1]) & (f[:, 1] != f[:, 2]) & (f[:, 0] != f[:, 2])
        area2 = np.linalg.norm(np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]]), axis=1)
        f = f[distinct & (area2 > 0.0)]
    return Mesh(v, f)


def load_mesh(path: str = BUNNY, height: float = 0.10, yaw_deg: float = 0.0, y_up: bool = True) -> Mesh:
    """Mesh scaled to `height` [m], Z-up, yawed, centred in xy with its lowest point at z = 0
    (shapes.py:20-32; the same float64 operations in the same order, so the same vertices)."""
    m = read_obj(path)
    v = m.vertices.copy()
    if y_up:                                        # Y-up file -> Z-up: (x, y, z) -> (x, -z, y)
        v = np.stack([v[:, 0], -v[:, 2], v[:, 1]], 1)
    c, s = np.cos(np.deg2rad(yaw_deg)), np.sin(np.deg2rad(yaw_deg))
    v = np.stack([c * v[:, 0] - s * v[:, 1], s * v[:, 0] + c * v[:, 1], v[:, 2]], 1)
    v -= v.min(0)
    v *= height / v[:, 2].max()
    v[:, :2] -= v.max(0)[:2] / 2
    return Mesh(v, m.faces)


def inside(mesh: Mesh, points: np.ndarray, device="cuda:0", threshold: float = 0.5, with_winding: bool = False):
    """Flags (1 inside) of float32 points against the mesh on the GPU (hfm_mesh_inside), and the
    winding numbers if asked."""
    from .solver import parse_device
    tri = mesh.triangles()
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    flags = np.zeros(len(pts), np.int32)
    wn = np.zeros(len(pts), np.float32) if with_winding else None
    lib = _abi.load()
    _abi.check(lib.hfm_mesh_inside(tri.ctypes.data_as(C.c_void_p), len(tri), pts.ctypes.data_as(C.c_void_p), len(pts),
                                   float(threshold), flags.ctypes.data_as(C.c_void_p),
                                   wn.ctypes.data_as(C.c_void_p) if wn is not None else None,
                                   parse_device(device)), "hfm_mesh_inside")
    return (flags, wn) if with_winding else flags


def lattice(mesh: Mesh, spacing: float, jitter: float, rng: np.random.Generator) -> np.ndarray:
    """The candidate points of sample_interior (shapes.py:36-39): a lattice of `spacing` over the
    mesh's box grown by one spacing, each point moved by U(-jitter, jitter)·spacing per axis."""
    lo, hi = mesh.vertices.min(0) - spacing, mesh.vertices.max(0) + spacing
    axes = [np.arange(lo[d] + spacing / 2, hi[d], spacing) for d in range(3)]
    P = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
    return P + rng.uniform(-jitter, jitter, P.shape) * spacing


def sample_interior(mesh: Mesh, spacing: float, jitter: float, rng: np.random.Generator, device) -> np.ndarray:
    """Candidate points inside the mesh (shapes.py:35-44).  The reference's query also asks for a
    closest point within the box's extent, which every candidate has, so only the winding number
    decides."""
    P = lattice(mesh, spacing, jitter, rng)
    keep = inside(mesh, P.astype(np.float32), device).astype(bool)
    return P[keep].astype(np.float32)


def mesh_particles(path: str = BUNNY, height: float = 0.10, spacing: float = 2.5e-3, L: float = 0.48, seed: int = 0,
                   yaw_deg: float = 0.0, device: str = "cuda:0") -> np.ndarray:
    """Particles filling the mesh, standing on the floor at the centre of a box of edge L
    (shapes.py:47-51)."""
    P = sample_interior(load_mesh(path, height, yaw_deg), spacing, 0.2, np.random.default_rng(seed), device)
    return (P + np.array([L / 2, L / 2, spacing / 2], np.float32)).astype(np.float32)
