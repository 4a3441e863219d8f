"""The one-launch pressure PCG (HFM_FLAG_FUSED_CG: warp's Jacobi PCG in the single-reduction
Chronopoulos-Gear form) against the fp32 oracle, which restates warp's two-pass PCG.  The two have the
same iterates in exact arithmetic and differ only in rounding (p.Ap and A p by recurrence), so the
one-launch form is held to the same bar as the default form: iteration counts within one, fields
within relative L2 1e-5 after a substep, 1e-4 after 60."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import rel_l2
from heatflow_mpm import GridConfig, HeatflowBatch, Heater, HeatflowMPM, coupled_scene, plume_scene

pytestmark = pytest.mark.gpu
FIELDS = ["T", "U", "V", "W", "P", "RHS"]


def _fused(kw, env=None):
    env = env or {}
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return HeatflowMPM(**kw, cg="fused")
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _check_one_substep(kw, dt, env=None, tol=1e-5):
    sim, orc = _fused(kw, env), O.OracleSim(**kw)
    sim.substep(dt)
    orc.substep(dt)
    st, so = sim.stats(), orc.stats()
    print("fused", st["p_iters_exact"], st["h_iters_exact"], "oracle", so["p_iters_exact"], so["h_iters_exact"],
          flush=True)
    assert abs(st["p_iters_exact"] - so["p_iters_exact"]) <= 1
    assert abs(st["h_iters_exact"] - so["h_iters_exact"]) <= 1
    if abs(st["p_iters_exact"] - so["p_iters_exact"]) == 0:
        assert st["p_converged"] == so["p_converged"]
    for f in FIELDS:
        err = rel_l2(sim._scene.get_field(f), orc.field(f))
        print(f, "rel_l2", err, flush=True)
        assert err <= tol, f
    return sim


def test_fused_plume_one_substep_matches_oracle(gpu):
    kw, dt, _ = plume_scene()
    _check_one_substep(kw, dt)


def test_fused_coupled_one_substep_matches_oracle(gpu):
    kw, dt, _ = coupled_scene(96)
    _check_one_substep(kw, dt)          # the first substep from rest runs to the 400-iteration cap


def test_fused_ragged_tiles_and_four_plane_blocks(gpu):
    """n = 40 (tiles cut by the box on every axis) with 8- and with 4-plane blocks."""
    kw, dt, _ = coupled_scene(40)
    _check_one_substep(kw, dt)
    _check_one_substep(kw, dt, env={"HFM_DIR_ICHUNK": "4"})


def test_fused_plume_many_substeps_track_the_oracle(gpu):
    kw, dt, _ = plume_scene()
    sim, orc = _fused(kw), O.OracleSim(**kw)
    worst = 0
    for _ in range(60):
        sim.substep(dt)
        orc.substep(dt)
        st, so = sim.stats(), orc.stats()
        worst = max(worst, abs(st["p_iters_exact"] - so["p_iters_exact"]))
    print("worst CG count difference", worst, flush=True)
    assert worst <= 1
    for f in ("T", "U", "V", "W", "P"):
        err = rel_l2(sim._scene.get_field(f), orc.field(f))
        print(f, "rel_l2 after 60 substeps", err, flush=True)
        assert err <= 1e-4, f


def test_fused_is_deterministic_and_the_same_in_a_batch(gpu):
    kw, dt, _ = coupled_scene(48)
    other = dict(particles=np.zeros((0, 3), np.float32), heater=Heater(power=400.0, t_cap=120.0, fan=2.0),
                 grid=GridConfig(L=0.32, h=0.01))
    a = _fused(kw)
    for _ in range(3):
        a.substep(dt)
    b, c = _fused(kw), HeatflowMPM(**other)        # fused and two-pass scenes in one batch
    ref_c = HeatflowMPM(**other)
    with HeatflowBatch([b, c]) as batch:
        for _ in range(3):
            batch.substep(dt)
            ref_c.substep(dt)
        for f in ("T", "U", "V", "W", "P"):
            assert np.array_equal(a._scene.get_field(f), b._scene.get_field(f)), f
            assert np.array_equal(c._scene.get_field(f), ref_c._scene.get_field(f)), f
        assert a.stats()["p_iters_total"] == b.stats()["p_iters_total"]


def test_fused_cg_cap_is_reported(gpu):
    kw, dt, _ = plume_scene()
    kw = dict(kw, grid=GridConfig(L=0.48, h=0.01, cg_tol=1e-9, cg_max_iter=7))
    sim, orc = _fused(kw), O.OracleSim(**kw)
    sim.substep(dt)
    orc.substep(dt)
    st, so = sim.stats(), orc.stats()
    assert st["p_iters_exact"] == so["p_iters_exact"] and st["p_converged"] == so["p_converged"] == 0
    assert st["p_iters"] == so["p_iters"]
    assert rel_l2(sim._scene.get_field("P"), orc.field("P")) <= 1e-5
