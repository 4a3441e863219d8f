"""Mesh seeding on the GPU (shapes.py:13-51): hfm_mesh_inside against the oracle's float64
winding number, and mesh_particles on meshes whose inside is known."""
import time

import numpy as np
import pytest

from _meshes import icosphere, triangles, write_obj
from heatflow_mpm import HeatflowMPM, Heater
from heatflow_mpm.shapes import Mesh, inside, lattice, load_mesh, mesh_particles
import winding as W

pytestmark = pytest.mark.gpu


def test_winding_number_matches_the_oracle(gpu):
    """A 5120-triangle blob (the bunny has 4968) and 6000 points in and around it: the same flags
    wherever the float64 winding number is not within 1e-3 of the threshold, fp32 winding numbers
    within 2e-4 of it."""
    v, f = icosphere(4, 0.05, (0.1, 0.2, 0.05), bump=0.25)
    mesh = Mesh(v, f)
    rng = np.random.default_rng(3)
    lo, hi = v.min(0) - 0.01, v.max(0) + 0.01
    pts = rng.uniform(lo, hi, (6000, 3)).astype(np.float32)
    t0 = time.perf_counter()
    flags, wn = inside(mesh, pts, with_winding=True)
    print(f"hfm_mesh_inside: {len(f)} triangles x {len(pts)} points in {1e3 * (time.perf_counter() - t0):.1f} ms "
          f"(incl. copies)", flush=True)
    big = rng.uniform(lo, hi, (50000, 3)).astype(np.float32)     # about the bunny's candidate count
    inside(mesh, big)
    t0 = time.perf_counter()
    inside(mesh, big)
    print(f"hfm_mesh_inside: {len(f)} triangles x {len(big)} points in {1e3 * (time.perf_counter() - t0):.1f} ms "
          f"(second call, incl. copies)", flush=True)
    w64 = W.winding_number(triangles(v, f), pts.astype(np.float64))
    clear = np.abs(w64 - 0.5) > 1e-3
    assert np.array_equal(flags[clear], (w64[clear] > 0.5).astype(np.int32))
    assert np.max(np.abs(wn - w64)) < 2e-4
    assert 0.2 < flags.mean() < 0.8


def test_mesh_particles_fill_a_sphere(gpu, tmp_path):
    """A sphere 10 cm high at 4 mm spacing: every particle inside the faceted sphere, the count
    near the volume over spacing^3, and the selection equal to the oracle's on the same
    candidates."""
    v, f = icosphere(3, 1.0)
    path = str(tmp_path / "sphere.obj")
    write_obj(path, v, f, y_up_from_z=True)
    s = 4e-3
    P = mesh_particles(path, height=0.10, spacing=s, L=0.48, seed=0)
    centre = np.array([0.24, 0.24, 0.05 + s / 2])
    r = np.linalg.norm(P - centre, axis=1)
    assert r.max() < 0.05 + 1e-6
    n_expect = 4.0 / 3.0 * np.pi * 0.05 ** 3 / s ** 3
    assert abs(len(P) / n_expect - 1.0) < 0.05, (len(P), n_expect)
    m = load_mesh(path, height=0.10)
    cand = lattice(m, s, 0.2, np.random.default_rng(0))
    w64 = W.winding_number(triangles(m.vertices, m.faces), cand.astype(np.float32).astype(np.float64))
    ref = cand[w64 > 0.5].astype(np.float32) + np.array([0.24, 0.24, s / 2], np.float32)
    clear = np.abs(w64 - 0.5) > 1e-3
    assert np.abs(len(P) - len(ref)) <= np.count_nonzero(~clear)
    if np.all(clear):
        assert np.array_equal(P, ref.astype(np.float32))


def test_seeded_scene_steps(gpu, tmp_path):
    v, f = icosphere(3, 1.0)
    path = str(tmp_path / "ball.obj")
    write_obj(path, v, f, y_up_from_z=True)
    P = mesh_particles(path, height=0.06, spacing=3e-3)
    sim = HeatflowMPM(P, heater=Heater(power=400.0, fan=2.0))
    for _ in range(3):
        sim.substep(2e-3)
    assert np.all(np.isfinite(sim.grid_T())) and np.all(np.isfinite(sim.temperatures()))
    assert abs(sim.positions()[:, 2].max() - 0.06) < 0.005
