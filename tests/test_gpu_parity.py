"""GPU parity of the whole substep against the fp32 oracle, through the C ABI.

The oracle restates the reference's kernels with the same fp32 operation order and the library
is compiled without FMA contraction, so without particles (no summation-order freedom) the GPU
fields match the oracle bit for bit up to the last-bit rounding of the fp64-accumulated dot
products; with particles the deterministic P2G sums in bin order instead of particle order.
Tolerances are written per test.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2
from heatflow_mpm import GridConfig, Heater, HeatflowMPM, coupled_scene, plume_scene

pytestmark = pytest.mark.gpu

FIELDS = ["T", "T2", "TW", "U", "V", "W", "P", "RHS", "ALPHA_S", "KAPPA", "PDIAG", "HRHS", "HDIAG"]


def _pair(kw):
    sim = HeatflowMPM(**kw)
    orc = O.OracleSim(**kw)
    return sim, orc


def test_device_visible(gpu):
    assert gpu >= 1


def test_plume_one_substep_matches_oracle(gpu):
    kw, dt, _ = plume_scene()
    sim, orc = _pair(kw)
    sim.substep(dt)
    orc.substep(dt)
    st, so = sim.stats(), orc.stats()
    print("gpu", st, "\noracle", so, flush=True)
    assert st["h_iters_exact"] == so["h_iters_exact"]
    assert abs(st["p_iters_exact"] - so["p_iters_exact"]) <= 1
    assert st["p_iters"] == so["p_iters"] and st["h_iters"] == so["h_iters"]
    for f in FIELDS:
        g, o = sim._scene.get_field(f), orc.field(f)
        err = rel_l2(g, o)
        print(f, "rel_l2", err, "max|d|", float(np.abs(g - o).max()), flush=True)
        assert err <= 1e-5, f


def test_coupled_one_substep_matches_oracle(gpu):
    kw, dt, _ = coupled_scene(96)
    sim, orc = _pair(kw)
    sim.substep(dt)
    orc.substep(dt)
    st, so = sim.stats(), orc.stats()
    print("gpu", st, "\noracle", so, flush=True)
    assert abs(st["p_iters_exact"] - so["p_iters_exact"]) <= 1
    assert abs(st["h_iters_exact"] - so["h_iters_exact"]) <= 1
    for f in FIELDS + ["ALPHA_P", "TSUM", "QX", "QY", "QZ", "KX", "KY", "KZ", "MX", "MY", "MZ"]:
        g, o = sim._scene.get_field(f), orc.field(f)
        err = rel_l2(g, o)
        print(f, "rel_l2", err, flush=True)
        assert err <= 1e-5, f
    tg, to = sim.temperatures(), orc.temperatures()
    assert rel_l2(tg, to) <= 1e-5


def test_plume_many_substeps_track_the_oracle(gpu):
    """Plume scene (n=48, PCG tol 1e-6): 60 substeps in lockstep; CG counts within +-1 at every
    substep, T and u within relative L2 1e-4 at the end (1e-3 is the bar after 500 substeps)."""
    kw, dt, _ = plume_scene()
    sim, orc = _pair(kw)
    worst = 0
    for step in range(60):
        sim.substep(dt)
        orc.substep(dt)
        st, so = sim.stats(), orc.stats()
        worst = max(worst, abs(st["p_iters_exact"] - so["p_iters_exact"]), abs(st["h_iters_exact"] - so["h_iters_exact"]))
    print("worst CG count difference", worst, flush=True)
    assert worst <= 1
    for f in ("T", "U", "V", "W", "P"):
        err = rel_l2(sim._scene.get_field(f), orc.field(f))
        print(f, "rel_l2 after 60 substeps", err, flush=True)
        assert err <= 1e-4, f


def test_coupled_few_substeps_track_the_oracle(gpu):
    kw, dt, _ = coupled_scene(96)
    sim, orc = _pair(kw)
    for step in range(3):
        sim.substep(dt)
        orc.substep(dt)
        st, so = sim.stats(), orc.stats()
        assert abs(st["p_iters_exact"] - so["p_iters_exact"]) <= 1
    for f in ("T", "U", "V", "W", "P"):
        err = rel_l2(sim._scene.get_field(f), orc.field(f))
        print(f, "rel_l2 after 3 substeps", err, flush=True)
        assert err <= 1e-5, f
    assert rel_l2(sim.temperatures(), orc.temperatures()) <= 1e-5
    assert rel_l2(sim.melt_fraction(), orc.melt_fraction()) <= 1e-5 or np.abs(sim.melt_fraction()).max() == 0


def test_substep_is_deterministic(gpu):
    """No float atomics anywhere: two runs of the coupled scene are bit-identical."""
    kw, dt, _ = coupled_scene(64)
    out = []
    for _ in range(2):
        sim = HeatflowMPM(**kw)
        for _ in range(5):
            sim.substep(dt)
        out.append([sim._scene.get_field(f) for f in ("T", "U", "V", "W", "P", "ALPHA_S", "TSUM")] + [sim.temperatures()])
        sim.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_particles_outside_the_box_are_clamped_like_the_reference(gpu):
    """fields.py:39-48 clamps the lattice indices: material outside the box lands on the edge nodes."""
    rng = np.random.default_rng(4)
    P = np.concatenate([rng.uniform(-0.02, 0.0, (50, 3)), rng.uniform(0.24, 0.26, (50, 3)),
                        rng.uniform(0.48, 0.50, (50, 3))]).astype(np.float32)
    kw = dict(particles=P, heater=Heater(power=300.0, fan=1.0), grid=GridConfig(L=0.32, h=0.01, cg_tol=1e-6))
    sim, orc = _pair(kw)
    sim.substep(2e-3)
    orc.substep(2e-3)
    for f in ("ALPHA_S", "ALPHA_P", "TSUM", "KX", "KY", "KZ", "T", "U", "W"):
        assert rel_l2(sim._scene.get_field(f), orc.field(f)) <= 1e-5, f
    assert rel_l2(sim.temperatures(), orc.temperatures()) <= 1e-5


def test_non_finite_input_is_rejected(gpu):
    kw, dt, _ = plume_scene()
    sim = HeatflowMPM(**kw)
    n = sim.grid.n
    vals = np.full((n, n, n), 25.0)
    vals[3, 3, 3] = np.nan
    with pytest.raises(RuntimeError, match="non-finite"):
        sim.heat.set_fixed(np.ones((n, n, n), bool), vals)
    with pytest.raises(RuntimeError, match="bad argument"):
        sim.substep(-1.0)
    sim.substep(dt)                                 # the scene is still usable
    assert np.isfinite(sim.grid_T()).all()


def test_cg_cap_is_reported_not_raised(gpu):
    """A solve that hits its iteration cap is a statistic (grid.py:248 discards the residual)."""
    kw, dt, _ = plume_scene()
    kw["grid"].cg_max_iter = 4
    sim = HeatflowMPM(**kw)
    for _ in range(3):
        sim.substep(dt)
    st = sim.stats()
    assert st["p_iters_exact"] == 4 and st["p_iters"] == 4 and st["p_converged"] == 0



@pytest.mark.parametrize("layout", [dict(dir_planes=-1), dict(dir_planes=4), dict(dir_planes=8),
                                    dict(dir_planes=16), dict(dir_planes=8, cg="fused")],
                         ids=["untiled", "4 planes", "8 planes", "16 planes", "fused on 8 planes"])
def test_layouts_match_the_oracle(gpu, layout):
    """Every layout the autotuner may pick gives the oracle's substep (n = 40: tiles cut by the box)."""
    kw, dt, _ = coupled_scene(40)
    sim = HeatflowMPM(**kw, layout=layout)
    assert sim.layout["dir_planes"] == layout["dir_planes"]
    orc = O.OracleSim(**kw)
    sim.substep(dt)
    orc.substep(dt)
    assert abs(sim.stats()["p_iters_exact"] - orc.stats()["p_iters_exact"]) <= 1
    for f in ("T", "U", "V", "W", "P"):
        assert rel_l2(sim._scene.get_field(f), orc.field(f)) <= 1e-5, f
