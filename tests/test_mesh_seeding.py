"""Mesh seeding on the host (shapes.py:1-51): the oracle's winding number pinned by closed forms,
the OBJ reader, the mesh transform and the candidate lattice; the C entry point's argument
checks (no GPU needed)."""
import ctypes as C
import os

import numpy as np
import pytest

from _meshes import cube, flip, icosphere, triangles, write_obj
from heatflow_mpm import _abi
from heatflow_mpm.shapes import Mesh, lattice, load_mesh, read_obj
import winding as W


def test_octant_triangle_subtends_an_eighth_of_the_sphere():
    tri = np.array([[1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0]])
    assert W.solid_angles(tri, np.zeros((1, 3)))[0, 0] == pytest.approx(np.pi / 2, rel=1e-14)
    # seen from the far side of the plane the sign turns
    assert W.solid_angles(tri, np.ones((1, 3)))[0, 0] < 0


def test_cube_inside_outside_and_orientation():
    tri = triangles(*cube())
    rng = np.random.default_rng(0)
    pin = rng.uniform(-0.95, 0.95, (200, 3))
    pout = rng.uniform(-3, 3, (400, 3))
    pout = pout[np.abs(pout).max(1) > 1.05]
    assert np.allclose(W.winding_number(tri, pin), 1.0, atol=1e-12)
    assert np.allclose(W.winding_number(tri, pout), 0.0, atol=1e-12)
    v, f = cube()
    assert np.allclose(W.winding_number(triangles(v, flip(f)), pin), -1.0, atol=1e-12)


def test_sphere_union_and_cavity():
    rng = np.random.default_rng(1)
    p = rng.uniform(-2.5, 2.5, (3000, 3))
    r = np.linalg.norm(p, axis=1)
    v1, f1 = icosphere(3, 2.0)
    v2, f2 = icosphere(3, 1.0)
    shell = np.concatenate([triangles(v1, f1), triangles(v2, flip(f2))])
    w = W.winding_number(shell, p)
    # the faceted spheres lie between 0.98 R and R; away from both, w is 1 in the shell, 0 elsewhere
    clear = (np.abs(r - 2.0) > 0.05) & (np.abs(r - 1.0) > 0.05)
    expect = ((r < 1.95) & (r > 1.05)).astype(float)
    assert np.allclose(w[clear], expect[clear], atol=1e-9)
    # two disjoint cubes: the union
    a, b = cube((-2, -2, -2), (-1, -1, -1)), cube((1, 1, 1), (2, 2, 2))
    two = np.concatenate([triangles(*a), triangles(*b)])
    q = np.array([[-1.5, -1.5, -1.5], [1.5, 1.5, 1.5], [0, 0, 0]])
    assert np.allclose(W.winding_number(two, q), [1, 1, 0], atol=1e-12)


def test_read_obj_triangulates_and_drops_degenerate_faces(tmp_path):
    path = tmp_path / "m.obj"
    path.write_text("# quad, a fan polygon, negative indices, v/vt/vn, a degenerate face\n"
                    "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 2 0 0\n"
                    "vt 0 0\nvn 0 0 1\n"
                    "f 1/1/1 2/1/1 3/1/1 4/1/1\n"
                    "f -5 -4 -1\n"
                    "f 1 2 2\n")
    m = read_obj(str(path))
    assert m.vertices.shape == (5, 3)
    assert m.faces.tolist() == [[0, 1, 2], [0, 2, 3]]   # (1, 2, 5) is colinear: zero area, dropped
    with pytest.raises(FileNotFoundError, match="no mesh is bundled"):
        read_obj(str(tmp_path / "missing.obj"))


def test_load_mesh_transform(tmp_path):
    """Y-up file to Z-up, yaw, floor at z = 0, scaled to the height, centred in xy (shapes.py:20-32)."""
    v, f = cube((0.0, 0.0, 0.0), (2.0, 1.0, 4.0))        # Z-up box 2 x 1 x 4
    path = str(tmp_path / "box.obj")
    write_obj(path, v, f, y_up_from_z=True)
    m = load_mesh(path, height=0.1)
    assert m.vertices[:, 2].min() == 0.0 and m.vertices[:, 2].max() == pytest.approx(0.1, rel=1e-15)
    ext = m.vertices.max(0) - m.vertices.min(0)
    assert np.allclose(ext, [0.05, 0.025, 0.1])
    assert np.allclose(m.vertices.max(0)[:2], -m.vertices.min(0)[:2])     # centred in x and y
    m90 = load_mesh(path, height=0.1, yaw_deg=90.0)                       # yaw swaps the x and y extents
    assert np.allclose(m90.vertices.max(0) - m90.vertices.min(0), [0.025, 0.05, 0.1])
    # the inside test sees the same closed box after the transform
    tri = triangles(m.vertices, m.faces)
    assert W.winding_number(tri, np.array([[0.0, 0.0, 0.05]]))[0] == pytest.approx(1.0, abs=1e-12)


def test_lattice_matches_the_reference_construction():
    m = Mesh(*cube((0.0, 0.0, 0.0), (0.01, 0.02, 0.03)))
    s = 2.5e-3
    P = lattice(m, s, 0.2, np.random.default_rng(0))
    lo, hi = m.vertices.min(0) - s, m.vertices.max(0) + s
    counts = [len(np.arange(lo[d] + s / 2, hi[d], s)) for d in range(3)]
    assert len(P) == int(np.prod(counts))
    base = np.stack(np.meshgrid(*[np.arange(lo[d] + s / 2, hi[d], s) for d in range(3)], indexing="ij"), -1)
    jit = (P - base.reshape(-1, 3)) / s
    assert np.all(np.abs(jit) <= 0.2)
    assert np.array_equal(P, lattice(m, s, 0.2, np.random.default_rng(0)))


def test_mesh_inside_argument_checks():
    lib = _abi.load()
    flags = np.zeros(4, np.int32)
    pts = np.zeros((4, 3), np.float32)
    tri = triangles(*cube()).astype(np.float32)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    assert lib.hfm_mesh_inside(vp(tri), -1, vp(pts), 4, 0.5, vp(flags), None, 0) == -1
    assert lib.hfm_mesh_inside(vp(tri), len(tri), None, 4, 0.5, vp(flags), None, 0) == -1
    assert lib.hfm_mesh_inside(vp(tri), len(tri), vp(pts), 0, 0.5, None, None, 0) == 0
    bad = pts.copy()
    bad[1, 2] = np.nan
    assert lib.hfm_mesh_inside(vp(tri), len(tri), vp(bad), 4, 0.5, vp(flags), None, 0) == -4
