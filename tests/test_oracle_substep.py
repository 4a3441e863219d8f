"""Pins the oracle's substep stages with closed forms, conservation laws and the reference's own
acceptance test (CPU only)."""
import math

import numpy as np
import pytest

import oracle as O
from conftest import rel_l2
from heatflow_mpm import GridConfig, Heater, Material, ellipsoid_particles

NONE = np.zeros((0, 3), np.float32)


def _prep(sim, dt):
    for st in ("CLEAR", "ADD_STATIC", "BLEND", "FIXED_TW"):
        sim.stage(st, dt)


def test_source_injects_the_heater_power():
    """heat.py:154-158 with rate from heater.py:52: sum dT rho c_p h^3 = P dt while no cell is capped."""
    dt = 2e-3
    sim = O.OracleSim(NONE, heater=Heater(power=400.0, t_cap=120.0), grid=GridConfig(L=0.48, h=0.01),
                      precision="f64")
    _prep(sim, dt)
    before = sim.field("TW")
    sim.stage("SOURCE", dt)
    after = sim.field("TW")
    src = sim.inputs["source"].astype(bool)
    assert np.all(after[~src] == before[~src])
    energy = (after - before)[src].sum() * 1.2 * 1005.0 * 0.01 ** 3
    assert energy == pytest.approx(400.0 * dt, rel=1e-9)


def test_source_cap():
    """T never exceeds t_cap through the source; a cell already above it keeps its value."""
    dt = 2e-3
    sim = O.OracleSim(NONE, heater=Heater(power=400.0, t_cap=30.0), grid=GridConfig(L=0.48, h=0.01), precision="f64")
    src = sim.inputs["source"].astype(bool) & ~sim.inputs["fixed"].astype(bool)   # walls reset first
    T = np.full((48, 48, 48), 25.0)
    T[src] = np.where(np.arange(src.sum()) % 2 == 0, 29.9, 31.0)
    sim.set_field("T", T)
    _prep(sim, dt)
    sim.stage("SOURCE", dt)
    out = sim.field("TW")[src]
    assert np.all(out[T[src] == 31.0] == 31.0)
    low = T[src] == 29.9
    assert np.allclose(out[low], np.minimum(29.9 + sim.rate * dt, 30.0))


def test_blend_takes_the_particle_temperature():
    """heat.py:21-28: cells covered by material (alpha > 0.05) take the particles' mean temperature."""
    P = ellipsoid_particles(semi_axes=(0.05, 0.05, 0.05), spacing=2.5e-3, jitter=0.0)
    sim = O.OracleSim(P, heater=Heater(on=False), grid=GridConfig(L=0.48, h=0.01), t_init=12.5, precision="f64")
    for st in ("CLEAR", "P2G", "ADD_STATIC", "BLEND"):
        sim.stage(st, 2e-3)
    alpha = sim.field("ALPHA_P")
    Tw = sim.field("TW")
    covered = alpha > 0.05
    assert covered.sum() > 100
    assert np.allclose(Tw[covered], 12.5, atol=1e-12)
    assert np.all(Tw[~covered] == 25.0)
    # P2G conserves the particle volume fraction: sum(alpha) h^3 = P * spacing^3 (all in the box)
    assert alpha.sum() * 0.01 ** 3 == pytest.approx(len(P) * 2.5e-3 ** 3, rel=1e-12)
    assert np.array_equal(sim.field("ALPHA_S"), alpha)            # no static solid with the heater off


def test_zero_velocity_advection_is_the_identity_and_solids_keep_their_value():
    dt = 2e-3
    sim = O.OracleSim(NONE, heater=Heater(), grid=GridConfig(L=0.48, h=0.01), precision="f64")
    rng = np.random.default_rng(0)
    sim.set_field("T", 25 + rng.standard_normal((48, 48, 48)))
    _prep(sim, dt)
    sim.stage("ADVECT_T", dt)
    assert np.allclose(sim.field("T2"), sim.field("TW"), atol=1e-9)


def test_uniform_flow_translates_a_linear_field():
    """Semi-Lagrangian advection is exact for a linear field under a uniform flow (away from walls)."""
    dt, n, h = 2e-3, 24, 0.02
    sim = O.OracleSim(NONE, heater=Heater(on=False), grid=GridConfig(L=n * h, h=h), precision="f64")
    U = (0.8, -0.5, 0.3)
    sim.set_field("U", np.full((n + 1, n, n), U[0]))
    sim.set_field("V", np.full((n, n + 1, n), U[1]))
    sim.set_field("W", np.full((n, n, n + 1), U[2]))
    g = (np.arange(n) + 0.5) * h
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    lin = lambda x, y, z: 20 + 3 * x - 2 * y + 5 * z
    sim.set_field("T", lin(X, Y, Z))
    _prep(sim, dt)
    sim.stage("ADVECT_T", dt)
    inner = (slice(2, -2),) * 3
    exp = lin(X - dt * U[0], Y - dt * U[1], Z - dt * U[2])
    assert np.allclose(sim.field("T2")[inner], exp[inner], atol=1e-11)


def test_buoyancy_of_a_uniform_excess_temperature():
    """grid.py:36-43: interior z faces gain dt g dT / (T_amb + 273.15); wall faces stay."""
    dt, n = 2e-3, 16
    sim = O.OracleSim(NONE, heater=Heater(on=False), grid=GridConfig(L=n * 0.01, h=0.01), precision="f64")
    sim.set_field("T", np.full((n, n, n), 35.0))
    w0 = np.random.default_rng(2).standard_normal((n, n, n + 1))
    sim.set_field("W", w0)
    sim.stage("BUOYANCY", dt)
    w = sim.field("W")
    d = dt * 9.81 * 10.0 / (25.0 + 273.15)
    assert np.allclose(w[:, :, 1:-1] - w0[:, :, 1:-1], d, rtol=1e-12)
    assert np.array_equal(w[:, :, [0, -1]], w0[:, :, [0, -1]])


def test_relaxation_limits():
    """grid.py:70-79: K = 0 leaves the velocity unchanged; a large K drives it to m / K."""
    dt, n = 2e-3, 8
    sim = O.OracleSim(NONE, heater=Heater(on=False), grid=GridConfig(L=n * 0.01, h=0.01), precision="f64")
    sim.stage("CLEAR", dt)
    sim.stage("FACE_FRACTION", dt)
    u0 = np.random.default_rng(3).standard_normal((n + 1, n, n))
    sim.set_field("U", u0)
    K = np.zeros((n + 1, n, n))
    K[3] = 1e9
    m = np.zeros_like(K)
    m[3] = 2.0e9
    sim.set_field("KX", K)
    sim.set_field("MX", m)
    sim.stage("RELAX", dt)
    u = sim.field("U")
    assert np.allclose(u[3], 2.0, rtol=1e-6)
    mask = np.ones(n + 1, bool)
    mask[3] = False
    assert np.array_equal(u[mask], u0[mask])


def test_projection_makes_the_mixture_flow_divergence_free():
    """grid.py:235-251: after the projection div(a_f u + q) is at the PCG tolerance."""
    kw = dict(particles=NONE, heater=Heater(power=400.0, fan=2.0), grid=GridConfig(L=0.32, h=0.01, cg_tol=1e-10))
    sim = O.OracleSim(**kw, precision="f64")
    for _ in range(3):
        sim.substep(2e-3)
    div_max, speed = sim.diagnostics()
    assert speed > 0 and div_max < 1e-9
    st = sim.stats()
    assert st["p_converged"] == 1 and st["p_rr"] <= st["p_atol"] ** 2


def test_particle_update_closed_forms():
    """heat.py:161-174: with no grid change T_p stays; melt, yield and viscosity follow T_p."""
    P = np.array([[0.2, 0.2, 0.2], [0.25, 0.21, 0.22], [0.3, 0.3, 0.3]], np.float32)
    mat = Material(t_sol=28.0, t_liq=34.0, yield_solid=2e4, viscosity_solid=5.0, viscosity_liquid=30.0)
    sim = O.OracleSim(P, mat=mat, heater=Heater(on=False), grid=GridConfig(L=0.48, h=0.01, buoyancy=False),
                      t_init=31.0, precision="f64")
    sim.set_particle_temperatures(np.array([20.0, 31.0, 40.0], np.float32))
    sim.substep(2e-3)                            # uniform 25 C air, nothing moves: dT = 0
    T, f, y, mu = sim.particles()
    assert np.allclose(T, [20.0, 31.0, 40.0], atol=1e-9)
    assert np.allclose(f, [0.0, 0.5, 1.0], atol=1e-9)
    assert np.allclose(y, 2e4 * (1 - f)) and np.allclose(mu, 5.0 * (1 - f) + 30.0 * f)


def test_fp32_oracle_tracks_the_fp64_oracle():
    kw = dict(particles=ellipsoid_particles(spacing=5e-3, jitter=0.5e-3), mat=Material(t_sol=28, t_liq=34),
              heater=Heater(power=400.0, fan=2.0), grid=GridConfig(L=0.48, h=0.02, cg_tol=1e-6), spacing=5e-3)
    a, b = O.OracleSim(**kw, precision="f32"), O.OracleSim(**kw, precision="f64")
    for _ in range(5):
        a.substep(2e-3)
        b.substep(2e-3)
    assert rel_l2(a.field("T"), b.field("T")) < 1e-5
    for f in ("U", "V", "W"):
        assert rel_l2(a.field(f), b.field(f)) < 2e-3


def test_implicit_conduction_matches_erfc_early_time():
    """The reference's conduction test (tests/test_solver.py:21-30) at t = 0.2 s on the fp32 oracle:
    hot floor at 70 C, everything else free, kappa = 1e-3; the full 2 s run is a GPU test."""
    sim = O.OracleSim(NONE, heater=Heater(on=False), grid=GridConfig(L=0.30, buoyancy=False), kappa_air=1e-3)
    n = sim.n
    fixed = np.zeros((n, n, n), bool)
    fixed[:, :, 0] = True
    sim.set_fixed(fixed, np.where(fixed, 70.0, 0.0))
    t = 0.2
    sim.run(2e-3, int(round(t / 2e-3)))
    T = sim.grid_T()[n // 2, n // 2, :]
    z = np.maximum((np.arange(n) + 0.5) * sim.h - sim.h / 2, 0.0)
    ref = 25.0 + 45.0 * np.array([math.erfc(v / (2 * math.sqrt(1e-3 * t))) for v in z])
    assert np.abs(T - ref)[: n // 2].max() < 0.1
