"""GPU checks of the parts the whole-field parity tests see only indirectly, through the C ABI:

* cell binning and lattice indices bit for bit (fields.py:5-10) against numpy's float32 op sequence;
* the diagnostics air_velocity, divergence_max and speed_max (grid.py:156-162, 253-263) against the
  oracle on the same fields, and on an analytic field against its closed form;
* the PCG breakdown stop against the oracle (a heat system made indefinite through a negative
  material conductivity, an input the host passes; every kernel exits normally);
* the heater energy balance (heater.py:52, heat.py:154-158), summed in fp64 on the device.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2
from heatflow_mpm import GridConfig, Heater, HeatflowMPM, Material, coupled_scene, ellipsoid_particles, plume_scene

pytestmark = pytest.mark.gpu

OFFS = np.array([[0.5, 0.5, 0.5], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]], np.float32)


def _probe_positions(n, h, rng):
    """Inside, on lattice and half-lattice points, and outside the box on every side."""
    L = n * h
    parts = [rng.uniform(0.0, L, (3000, 3)),
             rng.integers(0, n + 1, (800, 3)) * h,                 # lattice points, the box faces included
             (rng.integers(0, n, (800, 3)) + 0.5) * h,             # half-lattice points
             rng.uniform(-3 * h, 0.0, (300, 3)), rng.uniform(L, L + 3 * h, (300, 3)),
             np.stack([rng.uniform(-h, L + h, 300), rng.uniform(0, L, 300), rng.uniform(-h, L + h, 300)], 1)]
    return np.concatenate(parts).astype(np.float32)


@pytest.mark.parametrize("n,L", [(48, 0.48), (96, 0.48), (40, 0.32)])
def test_particle_bins_and_indices_are_bit_exact(gpu, n, L):
    rng = np.random.default_rng(n)
    grid = GridConfig(L=L, h=L / n, cg_tol=1e-6)
    h32 = np.float32(grid.h)
    pos = _probe_positions(n, float(h32), rng)
    sim = HeatflowMPM(pos, heater=Heater(power=300.0), grid=grid)
    b, base, frac = sim._scene.particle_bins()
    q = pos / h32                                                  # float32 division, as fields.py:8
    for lat in range(4):
        g = q - OFFS[lat]                                          # np.float32(pos)/np.float32(h) - off
        fl = np.floor(g)
        assert np.array_equal(base[:, lat, :], fl.astype(np.int32)), lat
        assert np.array_equal(frac[:, lat, :].view(np.uint32), (g - fl).astype(np.float32).view(np.uint32)), lat
    cell = np.clip(np.floor(q).astype(np.int64), 0, n - 1)
    assert np.array_equal(b, ((cell[:, 0] * n + cell[:, 1]) * n + cell[:, 2]).astype(np.int32))
    # and the oracle's own base_index (fp32) on a sample of them
    for t in range(0, len(pos), 97):
        for lat in range(4):
            bi, f = O.base_index(pos[t].astype(np.float64), OFFS[lat].astype(np.float64), float(h32))
            assert np.array_equal(bi, base[t, lat]) and np.array_equal(np.float32(f), frac[t, lat]), (t, lat)
    # a substep after the query bins nothing twice: it matches the oracle
    orc = O.OracleSim(pos, heater=Heater(power=300.0), grid=grid)
    sim.substep(2e-3)
    orc.substep(2e-3)
    for f in ("ALPHA_S", "ALPHA_P", "TSUM", "KX", "T", "U"):
        assert rel_l2(sim._scene.get_field(f), orc.field(f)) <= 1e-5, f
    sim.close()


def test_diagnostics_match_the_oracle_on_the_same_fields(gpu):
    kw, dt, _ = coupled_scene(48)
    sim = HeatflowMPM(**kw)
    orc = O.OracleSim(**kw)
    for _ in range(3):
        sim.substep(dt)
        orc.substep(dt)
    # the oracle's own state: both diagnostics agree to the fields' parity
    dm_g, sm_g = sim.grid.divergence_max(), sim.grid.speed_max()
    dm_o, sm_o = orc.diagnostics()
    print("own states: divergence_max", dm_g, dm_o, "speed_max", sm_g, sm_o, flush=True)
    assert sm_g == pytest.approx(sm_o, rel=1e-4)
    assert dm_g < 1e-3 and dm_o < 1e-3
    assert rel_l2(sim.air_velocity(), orc.air_velocity()) <= 1e-5
    # the same inputs in both: every field the three functions read, copied from the device
    for f in ("U", "V", "W", "ALPHA_S", "QX", "QY", "QZ"):
        orc.set_field(f, sim._scene.get_field(f).astype(np.float64))
    dm_o, sm_o = orc.diagnostics()
    print("same fields: divergence_max", dm_g, dm_o, "speed_max", sm_g, sm_o, flush=True)
    assert sm_g == sm_o
    assert dm_g == pytest.approx(dm_o, rel=1e-6)
    cv = sim.air_velocity()
    assert np.array_equal(cv, orc.air_velocity().astype(np.float32))
    n = sim.grid.n
    u, v, w = (sim._scene.get_field(f) for f in ("U", "V", "W"))
    ref = np.stack([np.float32(0.5) * (u[:-1] + u[1:]), np.float32(0.5) * (v[:, :-1] + v[:, 1:]),
                    np.float32(0.5) * (w[:, :, :-1] + w[:, :, 1:])], -1).reshape(n ** 3, 3)
    assert np.array_equal(cv, ref)
    sim.close()


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_divergence_max_of_a_linear_field(gpu, axis):
    """Solid-free 8^3 box (the heater lies outside it), face velocity 0.25 x (face index) along one
    axis: divergence_max = 0.25 (n - 1) at the cell beside the far wall, speed_max = 0.25 n."""
    n = 8
    sim = HeatflowMPM(np.zeros((0, 3), np.float32), heater=Heater(power=0.0), grid=GridConfig(L=0.08, h=0.01))
    sim._scene.set_field("ALPHA_S", np.zeros((n, n, n), np.float32))
    shape = [n, n, n]
    shape[axis] = n + 1
    idx = [1, 1, 1]
    idx[axis] = n + 1
    field = np.broadcast_to((0.25 * np.arange(n + 1, dtype=np.float32)).reshape(idx), shape).copy()
    sim._scene.set_field(["U", "V", "W"][axis], field)
    assert sim.grid.divergence_max() == pytest.approx(0.25 * (n - 1), rel=1e-6)
    assert sim.grid.speed_max() == 0.25 * n
    sim.close()


@pytest.mark.parametrize("kappa", [-0.1, -0.02])
def test_heat_breakdown_stop_matches_the_oracle(gpu, kappa):
    """kappa_mat < 0 makes the heat system indefinite inside the material: the solve meets a p.Ap or
    r.z that is not positive and stops (converged = -1) on both sides, at the same iteration, with
    the state finite."""
    L, n = 0.48, 24
    parts = ellipsoid_particles(semi_axes=(0.06, 0.05, 0.06), spacing=5e-3, jitter=0.6e-3, seed=1, L=L)
    kw = dict(particles=parts, mat=Material(t_sol=28.0, t_liq=34.0, kappa=kappa),
              heater=Heater(power=400.0, t_cap=120.0, fan=2.0), grid=GridConfig(L=L, h=L / n, cg_tol=1e-6),
              spacing=5e-3)
    sim = HeatflowMPM(**kw)
    orc = O.OracleSim(**kw)
    for step in range(3):
        sim.substep(2e-3)
        orc.substep(2e-3)
        st, so = sim.stats(), orc.stats()
        print(step, "gpu", st["h_iters_exact"], st["h_converged"], "oracle", so["h_iters_exact"], so["h_converged"],
              flush=True)
        assert st["h_converged"] == -1 and so["h_converged"] == -1
        assert abs(st["h_iters_exact"] - so["h_iters_exact"]) <= 1
        assert np.isfinite(sim.grid_T()).all()
    assert rel_l2(sim.grid_T(), orc.field("T")) <= (1e-5 if kappa == -0.1 else 1e-3)
    sim.close()


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_source_energy_balance(gpu, precision):
    """sum of the source step's rises x rho c_p h^3 = P dt while no source cell is capped; the device's
    fp64 sum equals the oracle's (same fp32 rises, another summation order)."""
    kw, dt, _ = plume_scene()
    sim = HeatflowMPM(**kw, precision=precision)
    orc = O.OracleSim(**kw)
    for step in range(2):
        sim.substep(dt)
        orc.substep(dt)
        g, o = sim.stats()["source_dT_sum"], orc.stats()["source_dT_sum"]
        print(precision, "device", repr(g), "oracle", repr(o), flush=True)
        # the first substep starts from the same T on both sides; later ones from T that agree to the
        # last bits (fp32) or to the mixed form's accuracy
        assert g == pytest.approx(o, rel=1e-12 if step == 0 else (1e-9 if precision == "fp32" else 1e-6))
        assert g * 1.2 * 1005.0 * 0.01 ** 3 == pytest.approx(400.0 * dt, rel=1e-5)
    sim.close()
