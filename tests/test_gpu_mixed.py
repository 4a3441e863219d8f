"""Mixed precision (HFM_PREC_MIXED): iterative refinement with FP16 inner PCG vectors, against the
fp64 oracle (the truth the mixed form is held to) and
against the fp32 form on the GPU.

Bars: T and u within relative L2 1e-3 of fp64 (the served form's tolerance); the pressure up to
its constant (the operator's null space) within 1e-3; divergence no worse than 1e-6 absolute;
iteration counts within 15 % of the fp32 solve (refinement restarts the Krylov space 2-3 times
per solve).  Batches: a mixed scene alone and in a batch give the same bits."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2
from heatflow_mpm import GridConfig, HeatflowBatch, Heater, HeatflowMPM, coupled_scene, plume_scene

pytestmark = pytest.mark.gpu
NONE = np.zeros((0, 3), np.float32)


def _u(sim_or_orc, get):
    return np.concatenate([np.asarray(get(sim_or_orc, f), np.float64).ravel() for f in ("U", "V", "W")])


def _gpu(sim, f):
    return sim._scene.get_field(f)


def _orc(orc, f):
    return orc.field(f)


def _centered(p):
    p = np.asarray(p, np.float64)
    return p - p.mean()


def test_mixed_plume_tracks_fp64(gpu):
    kw, dt, _ = plume_scene()
    K = 60
    mixed = HeatflowMPM(**kw, precision="mixed")
    plain = HeatflowMPM(**kw)
    its_m = its_p = 0
    for _ in range(K):
        mixed.substep(dt)
        plain.substep(dt)
        its_m += mixed.stats()["p_iters_exact"]
        its_p += plain.stats()["p_iters_exact"]
    st = mixed.stats()
    print("pressure iterations mixed", its_m, "fp32", its_p, "outer steps last substep", st["p_outer"],
          st["h_outer"], flush=True)
    f64 = O.OracleSim(**kw, precision="f64")
    f64.run(dt, K)
    eT = rel_l2(_gpu(mixed, "T"), _orc(f64, "T"))
    eU = rel_l2(_u(mixed, _gpu), _u(f64, _orc))
    eP = rel_l2(_centered(_gpu(mixed, "P")), _centered(_orc(f64, "P")))
    eT32 = rel_l2(_gpu(plain, "T"), _orc(f64, "T"))
    eU32 = rel_l2(_u(plain, _gpu), _u(f64, _orc))
    dm, _ = mixed._scene.diagnostics()
    dp, _ = plain._scene.diagnostics()
    print(f"vs fp64 after {K} substeps: mixed T {eT:.2e} u {eU:.2e} p {eP:.2e} div {dm:.2e}; "
          f"fp32 T {eT32:.2e} u {eU32:.2e} div {dp:.2e}", flush=True)
    assert eT <= 1e-3 and eU <= 1e-3 and eP <= 1e-3
    assert dm <= 1e-6
    assert st["p_converged"] in (1, 2) and st["h_converged"] in (1, 2) and st["p_outer"] >= 1 and st["h_outer"] >= 1
    assert its_m <= 1.15 * its_p


def test_mixed_coupled_substeps_track_fp64(gpu):
    kw, dt, _ = coupled_scene(64, particles_target=8000, seed=2)
    K = 3
    mixed = HeatflowMPM(**kw, precision="mixed")
    f64 = O.OracleSim(**kw, precision="f64")
    for _ in range(K):
        mixed.substep(dt)
        f64.substep(dt)
    eT = rel_l2(_gpu(mixed, "T"), _orc(f64, "T"))
    eU = rel_l2(_u(mixed, _gpu), _u(f64, _orc))
    eTp = rel_l2(mixed.temperatures(), f64.temperatures())
    print(f"coupled n=64 after {K} substeps vs fp64: T {eT:.2e} u {eU:.2e} particles {eTp:.2e}; "
          f"pressure its {mixed.stats()['p_iters_exact']} converged {mixed.stats()['p_converged']}", flush=True)
    assert eT <= 1e-3 and eU <= 1e-3 and eTp <= 1e-3


def test_mixed_alone_equals_mixed_in_a_batch(gpu):
    coupled40, dt, _ = coupled_scene(40)
    kws = [dict(particles=NONE, heater=Heater(power=400.0, t_cap=120.0, fan=2.0), grid=GridConfig(L=0.32, h=0.01)),
           coupled40,
           dict(particles=NONE, heater=Heater(power=300.0, t_cap=90.0),
                grid=GridConfig(L=0.48, h=0.01, cg_tol=1e-7, cg_max_iter=25))]
    alone = []
    for kw in kws:
        sim = HeatflowMPM(**kw, precision="mixed")
        for _ in range(3):
            sim.substep(dt)
        alone.append(({f: _gpu(sim, f) for f in ("T", "U", "V", "W", "P")}, sim.stats()))
        sim.close()
    sims = [HeatflowMPM(**kw, precision="mixed") for kw in kws]
    with HeatflowBatch(sims) as batch:
        for _ in range(3):
            batch.substep(dt)
        for sim, (fields, st) in zip(sims, alone):
            for f, v in fields.items():
                assert np.array_equal(_gpu(sim, f), v), f
            for k in ("p_iters_exact", "h_iters_exact", "p_outer", "h_outer", "p_converged", "p_rr", "p_atol"):
                assert sim.stats()[k] == st[k], k
    assert alone[2][1]["p_iters_exact"] <= 25


def test_mixed_iteration_cap(gpu):
    kw, dt, _ = plume_scene()
    kw = dict(kw, grid=GridConfig(L=0.48, h=0.01, cg_tol=1e-9, cg_max_iter=7))
    sim = HeatflowMPM(**kw, precision="mixed")
    sim.substep(dt)
    st = sim.stats()
    assert st["p_iters_exact"] == 7 and st["p_converged"] == 0 and st["p_iters"] == 7
    assert np.isfinite(st["p_rr"]) and np.all(np.isfinite(_gpu(sim, "P")))


def test_velocity_written_by_the_host_reaches_the_advection(gpu):
    """Fields copied from one mixed scene into another through set_field make the next substep
    identical."""
    kw, dt, _ = plume_scene()
    a = HeatflowMPM(**kw, precision="mixed")
    for _ in range(5):
        a.substep(dt)
    b = HeatflowMPM(**kw, precision="mixed")
    for f in ("T", "U", "V", "W", "P"):
        b._scene.set_field(f, _gpu(a, f))
    a.substep(dt)
    b.substep(dt)
    for f in ("T", "U", "V", "W", "P"):
        assert np.array_equal(_gpu(a, f), _gpu(b, f)), f


def test_batches_hold_one_precision(gpu):
    a = HeatflowMPM(NONE, heater=Heater(on=False), grid=GridConfig(L=0.16, h=0.01))
    b = HeatflowMPM(NONE, heater=Heater(on=False), grid=GridConfig(L=0.16, h=0.01), precision="mixed")
    with pytest.raises(RuntimeError, match="one precision"):
        HeatflowBatch([a, b])
