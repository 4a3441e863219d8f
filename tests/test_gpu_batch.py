"""Batches of scenes (hfm_batch_*): a scene stepped in a batch gives the same bits as the same
scene stepped alone, whatever else is in the batch (sizes, particles, heaters, tolerances,
iteration caps, pressure pass-A tilings).  Each scene's arithmetic and reduction trees depend
only on the scene, so the comparison is exact."""
import os

import numpy as np
import pytest

from heatflow_mpm import GridConfig, HeatflowBatch, Heater, HeatflowMPM, Material, coupled_scene, ellipsoid_particles

pytestmark = pytest.mark.gpu
NONE = np.zeros((0, 3), np.float32)
FIELDS = ["T", "T2", "TW", "U", "V", "W", "P", "RHS", "ALPHA_S"]
DT = 2e-3


def _mix():
    """Scene constructors of a heterogeneous batch: (name, kwargs, env)."""
    coupled40, _, _ = coupled_scene(40)
    coupled64, _, _ = coupled_scene(64, particles_target=8000, seed=3)
    return [
        ("plume32", dict(particles=NONE, heater=Heater(power=400.0, t_cap=120.0), grid=GridConfig(L=0.32, h=0.01)), {}),
        ("coupled40", coupled40, {}),
        ("fan48_cap", dict(particles=NONE, heater=Heater(power=300.0, t_cap=90.0, fan=2.0),
                           grid=GridConfig(L=0.48, h=0.01, cg_tol=1e-5, cg_max_iter=30)), {}),
        ("coupled64", coupled64, {}),
        ("conduction24", dict(particles=NONE, heater=Heater(on=False), grid=GridConfig(L=0.24, h=0.01, buoyancy=False),
                              kappa_air=1e-3), {}),
        ("flat36", dict(particles=ellipsoid_particles(semi_axes=(0.03, 0.03, 0.03), spacing=4e-3),
                        mat=Material(t_sol=22.0, t_liq=30.0), heater=Heater(power=500.0, t_cap=110.0, fan=1.0),
                        grid=GridConfig(L=0.36, h=0.01), spacing=4e-3, t_init=21.0), {"HFM_PCG_FLAT_DIR": "1"}),
        ("plume192", dict(particles=NONE, heater=Heater(power=400.0, t_cap=120.0), grid=GridConfig(L=0.48, h=0.0025)),
         {}),
    ]


def _make(kw, env):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return HeatflowMPM(**kw)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _state(sim):
    out = {f: sim._scene.get_field(f) for f in FIELDS}
    out["temp"] = sim.temperatures().copy()
    out["melt"] = sim.melt_fraction().copy()
    return out


def _assert_same(name, a, b, sa, sb):
    for k in a:
        assert np.array_equal(a[k], b[k]), f"{name}: {k} differs (max |d| {np.abs(a[k] - b[k]).max()})"
    for k in ("p_iters", "p_iters_exact", "h_iters", "h_iters_exact", "p_converged", "h_converged", "substeps",
              "p_iters_total", "h_iters_total"):
        assert sa[k] == sb[k], f"{name}: stats {k} {sa[k]} != {sb[k]}"
    for k in ("p_rr", "p_atol", "h_rr", "h_atol", "rhs_mean"):
        assert sa[k] == sb[k], f"{name}: stats {k} {sa[k]} != {sb[k]}"


def test_scene_alone_equals_scene_in_a_mixed_batch(gpu):
    mix = _mix()
    steps = 4
    alone = {}
    for name, kw, env in mix:
        sim = _make(kw, env)
        for _ in range(steps):
            sim.substep(DT)
        alone[name] = (_state(sim), sim.stats())
        print(name, "alone", sim.stats()["p_iters_exact"], sim.stats()["h_iters_exact"], flush=True)
        sim.close()
    sims = [_make(kw, env) for _, kw, env in mix]
    with HeatflowBatch(sims) as batch:
        for _ in range(steps):
            batch.substep(DT)
        for (name, _, _), sim in zip(mix, sims):
            st = sim.stats()
            print(name, "batch", st["p_iters_exact"], st["h_iters_exact"], flush=True)
            _assert_same(name, _state(sim), alone[name][0], st, alone[name][1])
    # the capped scene did hit its cap, so the batch mixed stopped and running solves
    assert alone["fan48_cap"][1]["p_converged"] == 0 or alone["fan48_cap"][1]["p_iters_exact"] <= 30


def test_interleaved_batch_and_solo_steps_stay_ordered(gpu):
    kw_a = dict(particles=NONE, heater=Heater(power=400.0, t_cap=120.0, fan=2.0), grid=GridConfig(L=0.32, h=0.01))
    kw_b, _, _ = coupled_scene(48)
    ref_a, ref_b = HeatflowMPM(**kw_a), HeatflowMPM(**kw_b)
    for _ in range(3):
        ref_a.substep(DT)
    for _ in range(2):
        ref_b.substep(DT)
    a, b = HeatflowMPM(**kw_a), HeatflowMPM(**kw_b)
    with HeatflowBatch([a, b]) as batch:
        batch.substep(DT)
        a.substep(DT)          # alone, on the scene's own stream, between two batch substeps
        batch.substep(DT)
        _assert_same("a", _state(a), _state(ref_a), a.stats(), ref_a.stats())
        _assert_same("b", _state(b), _state(ref_b), b.stats(), ref_b.stats())


def test_batch_run_and_timing(gpu):
    sims = [HeatflowMPM(NONE, heater=Heater(power=400.0, t_cap=120.0), grid=GridConfig(L=0.32, h=0.01))
            for _ in range(8)]
    with HeatflowBatch(sims) as batch:
        batch.set_kernel_timing(True)
        ms = batch.run(DT, 5)
        t_dir, n_dir = batch.kernel_time(0)
        t_upd, n_upd = batch.kernel_time(1)
        print("8 x n=32, 5 substeps:", ms, "ms; pass A", t_dir / max(n_dir, 1), "ms x", n_dir, flush=True)
        assert ms > 0 and n_dir > 0 and n_upd > 0 and t_dir > 0
    ref = [s._scene.get_field("T") for s in sims]
    for r in ref[1:]:
        assert np.array_equal(r, ref[0])       # identical scenes stay identical


def test_scene_held_by_a_batch_cannot_be_destroyed(gpu):
    sim = HeatflowMPM(NONE, heater=Heater(on=False), grid=GridConfig(L=0.16, h=0.01))
    batch = HeatflowBatch([sim])
    with pytest.raises(RuntimeError, match="bad state"):
        sim.close()
    batch.close()
    sim.close()


def test_batch_rejects_duplicates(gpu):
    sim = HeatflowMPM(NONE, heater=Heater(on=False), grid=GridConfig(L=0.16, h=0.01))
    with pytest.raises(RuntimeError, match="twice"):
        HeatflowBatch([sim, sim])
    sim.close()
