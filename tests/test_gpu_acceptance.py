"""The reference's own acceptance tests (tests/test_solver.py:21-54) on the GPU path, at their
full lengths and thresholds; the cold-solid test keeps its temperature/melt half with a static
particle body (its height half needs Newton's MPM, outside this path).  The plume and fan tests also
run with the multigrid-preconditioned pressure PCG and with limited MacCormack advection."""
import math

import numpy as np
import pytest

from heatflow_mpm import GridConfig, Heater, HeatflowMPM, Material, ellipsoid_particles

pytestmark = pytest.mark.gpu
NONE = np.zeros((0, 3), np.float32)


def run(sim, seconds, dt=2e-3):
    for _ in range(int(seconds / dt)):
        sim.substep(dt)


def test_implicit_conduction_matches_erfc(gpu):
    sim = HeatflowMPM(NONE, heater=Heater(on=False), grid=GridConfig(L=0.30, buoyancy=False), kappa_air=1e-3)
    n = sim.grid.n
    fixed = np.zeros((n, n, n), bool)
    fixed[:, :, 0] = True
    sim.heat.set_fixed(fixed, np.where(fixed, 70.0, 0.0))
    run(sim, 2.0)
    T = sim.grid_T()[n // 2, n // 2, :]
    z = np.maximum((np.arange(n) + 0.5) * sim.grid.h - sim.grid.h / 2, 0.0)
    ref = 25.0 + 45.0 * np.array([math.erfc(v / (2 * math.sqrt(1e-3 * 2.0))) for v in z])
    err = np.abs(T - ref)[: n // 2].max()
    print("erfc max error", err, flush=True)
    assert err < 0.1


@pytest.mark.parametrize("form", [{}, {"cg": "mg"}, {"advection": "maccormack"}], ids=["plain", "mg", "maccormack"])
def test_heater_plume_rises(gpu, form):
    sim = HeatflowMPM(NONE, heater=Heater(power=400.0, t_cap=120.0), **form)
    run(sim, 3.0)
    T = sim.grid_T()
    hc = sim.heater_geom["centre"]
    i, j = int(hc[0] / sim.grid.h), int(hc[1] / sim.grid.h)
    print("speed", sim.grid.speed_max(), "T above heater", T[i, j, sim.grid.n - 2], "div", sim.grid.divergence_max(),
          flush=True)
    assert np.isfinite(T).all() and sim.grid.speed_max() > 0.2
    assert T[i, j, sim.grid.n - 2] > sim.grid.cfg.t_amb + 3.0
    assert sim.grid.divergence_max() < 1e-3


@pytest.mark.parametrize("form", [{}, {"cg": "mg"}, {"advection": "maccormack"}], ids=["plain", "mg", "maccormack"])
def test_fan_heats_the_material_position(gpu, form):
    sim = HeatflowMPM(NONE, heater=Heater(power=400.0, t_cap=120.0, fan=2.0), **form)
    run(sim, 3.0)
    T = sim.grid_T()
    c = sim.grid.n // 2
    kb = int(0.05 / sim.grid.h)
    print("T at material position", T[c, c, kb], T[c, c, kb + 4], flush=True)
    assert max(T[c, c, kb], T[c, c, kb + 4]) > 40.0


def test_cold_solid_stays_solid(gpu):
    P = ellipsoid_particles(semi_axes=(0.04, 0.04, 0.05))
    sim = HeatflowMPM(P, heater=Heater(on=False), t_init=20.0)
    run(sim, 1.0)
    assert sim.temperatures().max() < sim.mat.t_sol and sim.melt_fraction().max() == 0.0
    assert np.all(sim.yield_stress() == np.float32(Material().yield_solid))
