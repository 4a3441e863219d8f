"""hfm_mesh_inside_noalloc (the inside test with its device staging in a fixed per-GPU array, no
hipMalloc) against hfm_mesh_inside through the C ABI: the same flags and winding numbers bit for bit,
on two calls in a row, the second one smaller; and the size limit."""
import ctypes as C

import numpy as np
import pytest

from _meshes import icosphere, triangles
from heatflow_mpm import _abi

pytestmark = pytest.mark.gpu


def _call(fn, tri, pts, threshold=0.5):
    np_ = len(pts)
    flags = np.full(np_, -7, np.int32)
    wn = np.full(np_, np.nan, np.float32)
    rc = fn(tri.ctypes.data_as(C.c_void_p), len(tri) // 9, pts.ctypes.data_as(C.c_void_p), np_, threshold,
            flags.ctypes.data_as(C.c_void_p), wn.ctypes.data_as(C.c_void_p), 0)
    return rc, flags, wn


def test_noalloc_matches_the_allocating_entry_point(gpu):
    lib = _abi.load()
    v, f = icosphere(4, 0.05, (0.1, 0.2, 0.05), bump=0.25)
    tri = np.ascontiguousarray(triangles(v, f), np.float32).reshape(-1)
    rng = np.random.default_rng(11)
    lo, hi = v.min(0) - 0.01, v.max(0) + 0.01
    for count in (40000, 3001):          # the second call smaller: stale results must not leak through
        pts = rng.uniform(lo, hi, (count, 3)).astype(np.float32)
        rc_a, fa, wa = _call(lib.hfm_mesh_inside, tri, pts)
        rc_b, fb, wb = _call(lib.hfm_mesh_inside_noalloc, tri, pts)
        assert rc_a == 0 and rc_b == 0, lib.hfm_last_error()
        assert np.array_equal(fa, fb)
        assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32))
        assert 0.1 < fb.mean() < 0.9
    # without winding numbers, and with no points at all
    pts = rng.uniform(lo, hi, (500, 3)).astype(np.float32)
    flags = np.zeros(500, np.int32)
    assert lib.hfm_mesh_inside_noalloc(tri.ctypes.data_as(C.c_void_p), len(tri) // 9, pts.ctypes.data_as(C.c_void_p),
                                       500, 0.5, flags.ctypes.data_as(C.c_void_p), None, 0) == 0
    assert np.array_equal(flags, _call(lib.hfm_mesh_inside, tri, pts)[1])
    assert lib.hfm_mesh_inside_noalloc(tri.ctypes.data_as(C.c_void_p), len(tri) // 9, None, 0, 0.5, None, None, 0) == 0


def test_noalloc_rejects_calls_beyond_its_scratch(gpu):
    lib = _abi.load()
    tri = np.zeros(9, np.float32)
    pts = np.zeros((500000, 3), np.float32)          # 5 x 500000 words > 2^21
    flags = np.zeros(len(pts), np.int32)
    rc = lib.hfm_mesh_inside_noalloc(tri.ctypes.data_as(C.c_void_p), 1, pts.ctypes.data_as(C.c_void_p), len(pts), 0.5,
                                     flags.ctypes.data_as(C.c_void_p), None, 0)
    assert rc == -1 and b"scratch" in lib.hfm_last_error()
