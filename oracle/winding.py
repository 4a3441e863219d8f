"""Exact generalized winding number of a triangle mesh, float64 numpy: the inside test of the
mesh seeding (shapes.py:13-17 `_inside`: inside where the winding number exceeds 0.5).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the heatflow_mpm package.

w(p) = (1 / 4 pi) * sum over triangles of the solid angle Omega the triangle subtends at p, with
tan(Omega / 2) = a . (b x c) / (|a||b||c| + (a . b)|c| + (b . c)|a| + (c . a)|b|) for a, b, c the
vertices relative to p (Van Oosterom and Strackee, IEEE Trans. Biomed. Eng. 30, 1983; the
generalized winding number of Jacobson, Kavan and Sorkine-Hornung, SIGGRAPH 2013).  Warp's
mesh_query_point_sign_winding_number evaluates an approximation of the same sum (fast winding
numbers with accuracy 2.0); warp is not installed here, so the approximation itself is not
reproduced: the tests pin this sum by closed forms (w = 1 inside and 0 outside closed meshes, the
octant triangle's solid angle pi / 2, orientation reversal, unions and cavities).
"""
from __future__ import annotations

import numpy as np


def solid_angles(tri: np.ndarray, p: np.ndarray) -> np.ndarray:
    """(P, F) solid angles of triangles tri (F, 9) at points p (P, 3), float64."""
    t = np.asarray(tri, dtype=np.float64).reshape(-1, 3, 3)
    q = np.asarray(p, dtype=np.float64).reshape(-1, 1, 3)
    a, b, c = t[None, :, 0] - q, t[None, :, 1] - q, t[None, :, 2] - q
    la, lb, lc = (np.linalg.norm(x, axis=-1) for x in (a, b, c))
    num = np.einsum("pfi,pfi->pf", a, np.cross(b, c))
    den = (la * lb * lc + np.einsum("pfi,pfi->pf", a, b) * lc + np.einsum("pfi,pfi->pf", b, c) * la
           + np.einsum("pfi,pfi->pf", c, a) * lb)
    return 2.0 * np.arctan2(num, den)


def winding_number(tri: np.ndarray, points: np.ndarray) -> np.ndarray:
    """Winding number of the mesh at every point (float64), in point chunks of about 2M
    point-triangle pairs (a few hundred MB of temporaries)."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    chunk = max(1, 2_000_000 // max(1, len(np.asarray(tri).reshape(-1, 9))))
    out = np.empty(len(pts))
    for s in range(0, len(pts), chunk):
        out[s:s + chunk] = solid_angles(tri, pts[s:s + chunk]).sum(axis=1) / (4.0 * np.pi)
    return out


def inside(tri: np.ndarray, points: np.ndarray, threshold: float = 0.5) -> np.ndarray:
    return (winding_number(tri, points) > threshold).astype(np.int32)
