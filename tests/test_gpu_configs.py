"""The batched, mixed-precision and served configurations at their own scale, on the GPU.

* C4, batched: the 64 uneven scenes (n in {32, 48, 64, 96, 128}, 100-800 W, fan 0-3 m/s, 2k-50k
  particles) with the autotuned batch layouts, stepped in one batch for the configuration's 200
  substeps; one scene of every size is stepped alone with the same layout and must give the same bits.
* C3, mixed precision: the coupled scene at n = 192 with about 50k particles, the configuration's 100
  substeps in the mixed form (FP16 PCG vectors inside iterative refinement) and in fp32: finite, projected (divergence_max small), mixed within 1e-3 of fp32.  The same scene at
  n = 48 against the fp64 oracle after 100 substeps: the error of T, u and div(u) of both forms, per
  result.
* C5, served: seeded Poisson streams over n in {48, 64, 96, 128, 192} in the mixed form with
  continuous batching, one with K 50-120 substeps (Jacobi-preconditioned refinement) and one over the
  configuration's K range 50-1000 (multigrid-preconditioned refinement, the served form); every result
  equals the request run alone, bit for bit, and the n = 48 results are held to 1e-3 against the fp64
  oracle.
"""
import time

import numpy as np
import pytest

import oracle as O
from conftest import rel_l2
from heatflow_mpm import HeatflowBatch, HeatflowMPM, batched_scenes, coupled_scene
from heatflow_mpm.autotune import lookup
from heatflow_mpm.serve import GpuBackend, Server, poisson_requests, summarize

pytestmark = pytest.mark.gpu


def _u(sim_or_fields):
    if isinstance(sim_or_fields, dict):
        return np.concatenate([sim_or_fields[f].ravel() for f in ("U", "V", "W")])
    return np.concatenate([sim_or_fields._scene.get_field(f).ravel() for f in ("U", "V", "W")])


def _u_orc(orc):
    return np.concatenate([orc.field(f).ravel() for f in ("U", "V", "W")])


def test_c4_batched_64_scenes_with_autotuned_layouts(gpu):
    scenes, dt, _ = batched_scenes(64)
    steps = 200
    t0 = time.time()
    sims = [HeatflowMPM(**kw, layout=lookup(kw["grid"].n, "batch")) for _, kw in scenes]
    cells = sum(s.grid.n ** 3 for s in sims)
    print(f"64 scenes, {cells / 1e6:.1f} M cells, {sum(s.n for s in sims)} particles, set up in "
          f"{time.time() - t0:.0f} s", flush=True)
    with HeatflowBatch(sims) as batch:
        t0 = time.perf_counter()
        ms = batch.run(dt, steps)
        print(f"batched: {steps} substeps in {ms:.0f} ms GPU, {cells * steps / (ms * 1e-3):.3g} cell-substeps/s",
              flush=True)
    picked = {}
    for i, (_, kw) in enumerate(scenes):
        picked.setdefault(kw["grid"].n, i)
    assert sorted(picked) == [32, 48, 64, 96, 128]
    for n, i in sorted(picked.items()):
        kw = scenes[i][1]
        alone = HeatflowMPM(**kw, layout=lookup(n, "batch"))
        alone.run(dt, steps)
        a, b = sims[i], alone
        for f in ("T", "U", "V", "W", "P"):
            assert np.array_equal(a._scene.get_field(f), b._scene.get_field(f)), (n, f)
        assert np.array_equal(a.temperatures(), b.temperatures()), n
        assert a.stats()["p_iters_total"] == b.stats()["p_iters_total"], n
        print(f"n={n}: alone == in the batch ({a.stats()['p_iters_total']} pressure iterations)", flush=True)
        alone.close()
    for s in sims:
        assert np.isfinite(s.grid_T()).all()
        s.close()


def test_c3_mixed_precision_n192_50k_particles(gpu):
    kw, dt, _ = coupled_scene(192, particles_target=50000)
    kw["grid"].cg_max_iter = 2000
    steps = 100
    mixed = HeatflowMPM(**kw, precision="mixed")
    plain = HeatflowMPM(**kw)
    print("particles", mixed.n, flush=True)
    t0 = time.perf_counter()
    mixed.run(dt, steps)
    mixed.synchronize()
    tm = time.perf_counter() - t0
    t0 = time.perf_counter()
    plain.run(dt, steps)
    plain.synchronize()
    tp = time.perf_counter() - t0
    eT, eu = rel_l2(mixed.grid_T(), plain.grid_T()), rel_l2(_u(mixed), _u(plain))
    dm, dp = mixed.grid.divergence_max(), plain.grid.divergence_max()
    sm, sp = mixed.stats(), plain.stats()
    print(f"n=192 {steps} substeps: mixed {tm * 1e3 / steps:.1f} ms/substep ({sm['p_iters_total']} pressure its), "
          f"fp32 {tp * 1e3 / steps:.1f} ms/substep ({sp['p_iters_total']}); mixed vs fp32: T {eT:.2e}, u {eu:.2e}; "
          f"divergence_max mixed {dm:.2e}, fp32 {dp:.2e}", flush=True)
    assert np.isfinite(mixed.grid_T()).all() and np.isfinite(_u(mixed)).all()
    assert eT <= 1e-3 and eu <= 1e-3
    assert dm < 1e-3 and dp < 1e-3
    mixed.close()
    plain.close()


def test_c3_same_scene_at_n48_against_fp64(gpu):
    kw, dt, _ = coupled_scene(48, particles_target=50000)
    steps = 100
    f64 = O.OracleSim(**kw, precision="f64")
    sims = {"mixed": HeatflowMPM(**kw, precision="mixed"), "fp32": HeatflowMPM(**kw)}
    f64.run(dt, steps)
    T64, u64 = f64.field("T"), _u_orc(f64)
    d64, _ = f64.diagnostics()
    for name, sim in sims.items():
        sim.run(dt, steps)
        eT, eu = rel_l2(sim.grid_T(), T64), rel_l2(_u(sim), u64)
        dm = sim.grid.divergence_max()
        print(f"n=48, {sim.n} particles, {steps} substeps, {name} vs fp64 oracle: T {eT:.2e}, u {eu:.2e}; "
              f"divergence_max {dm:.2e} (fp64 {d64:.2e})", flush=True)
        assert eT <= 1e-3 and eu <= 1e-3 and dm < 1e-3
        sim.close()


def test_c5_served_mixed_stream_equals_alone(gpu):
    reqs = poisson_requests(50.0, 8, seed=3, k_range=(50, 120), cg_max_iter=2000)
    sizes = [q.kwargs["grid"].n for q in reqs]
    assert set(sizes) == {48, 64, 96, 128, 192}
    srv = Server(GpuBackend(precision="mixed"), keep_results=True, max_active=16)
    t0 = time.perf_counter()
    res = srv.run(reqs)
    span = time.perf_counter() - t0
    s = summarize(res, span)
    print(f"served (mixed, continuous batching): {s['cell_substeps_per_s']:.3g} cell-substeps/s, p50 {s['p50_s']:.2f} s, "
          f"p99 {s['p99_s']:.2f} s, most scenes in one batch {max(e[3] for e in srv.log)}", flush=True)
    assert len(res) == len(reqs)
    for r, q in zip(res, reqs):
        alone = HeatflowMPM(**q.kwargs, precision="mixed")
        alone.run(q.dt, q.substeps)
        assert np.array_equal(alone.grid_T(), r.fields["T"]), r.rid
        for f in ("U", "V", "W"):
            assert np.array_equal(alone._scene.get_field(f), r.fields[f]), (r.rid, f)
        assert np.array_equal(alone.temperatures(), r.fields["temperatures"]), r.rid
        msg = f"request {r.rid}: n={r.n} K={r.substeps} alone == served"
        if r.n == 48:
            orc = O.OracleSim(**q.kwargs, precision="f64")
            orc.run(q.dt, q.substeps)
            eT, eu = rel_l2(r.fields["T"], orc.field("T")), rel_l2(_u(r.fields), _u_orc(orc))
            dm = alone.grid.divergence_max()
            msg += f"; vs fp64 oracle T {eT:.2e}, u {eu:.2e}, divergence_max {dm:.2e}"
            assert eT <= 1e-3 and eu <= 1e-3 and dm < 1e-3
        print(msg, flush=True)
        alone.close()


def test_c5_served_stream_at_configuration_scale(gpu):
    """The configuration's stream: K drawn from 50-1000 substeps over all five sizes (seed 3: 8.1e9
    cell-substeps), served in the mixed form with the multigrid-preconditioned refinement and continuous
    batching; each result equals the request run alone, and both n = 48 requests (275 and 158 substeps)
    are within 1e-3 of the fp64 oracle."""
    reqs = poisson_requests(50.0, 8, seed=3, k_range=(50, 1000), cg_max_iter=2000)
    assert {q.kwargs["grid"].n for q in reqs} == {48, 64, 96, 128, 192}
    srv = Server(GpuBackend(precision="mixed", cg="mg"), keep_results=True, max_active=16)
    t0 = time.perf_counter()
    res = srv.run(reqs)
    span = time.perf_counter() - t0
    s = summarize(res, span)
    print(f"served (mixed, multigrid, continuous batching), K {min(q.substeps for q in reqs)}-"
          f"{max(q.substeps for q in reqs)}: {s['cell_substeps']:.3g} cell-substeps, "
          f"{s['cell_substeps_per_s']:.3g} cell-substeps/s, p50 {s['p50_s']:.2f} s, p99 {s['p99_s']:.2f} s",
          flush=True)
    assert len(res) == len(reqs)
    for r, q in zip(res, reqs):
        alone = HeatflowMPM(**q.kwargs, precision="mixed", cg="mg")
        alone.run(q.dt, q.substeps)
        assert np.array_equal(alone.grid_T(), r.fields["T"]), r.rid
        for f in ("U", "V", "W"):
            assert np.array_equal(alone._scene.get_field(f), r.fields[f]), (r.rid, f)
        assert np.array_equal(alone.temperatures(), r.fields["temperatures"]), r.rid
        msg = f"request {r.rid}: n={r.n} K={r.substeps} alone == served"
        if r.n == 48:
            orc = O.OracleSim(**q.kwargs, precision="f64")
            orc.run(q.dt, q.substeps)
            eT, eu = rel_l2(r.fields["T"], orc.field("T")), rel_l2(_u(r.fields), _u_orc(orc))
            dm = alone.grid.divergence_max()
            msg += f"; vs fp64 oracle T {eT:.2e}, u {eu:.2e}, divergence_max {dm:.2e}"
            assert eT <= 1e-3 and eu <= 1e-3 and dm < 1e-3
        print(msg, flush=True)
        alone.close()


def test_c4_mixed_form_batch_equals_alone(gpu):
    """The 64 batched scenes in the mixed form, one batch, against n = 48, 64 and 128 scenes stepped
    alone.  A batch this large runs each launch in several rounds of workgroups, so a workgroup of a
    scene can start after workgroup 0 of the same scene finished: every value a launch decides with
    must come from earlier launches (a refinement step once read its predecessor's residual norm
    from a slot that its own workgroup 0 rewrote, and the late workgroups then took the solve for
    stalled)."""
    scenes, dt, _ = batched_scenes(64)
    pick = [50, 9, 3, 0]
    assert [scenes[i][1]["grid"].n for i in pick] == [48, 48, 64, 128]
    alone = {i: HeatflowMPM(**scenes[i][1], precision="mixed") for i in pick}
    sims = [HeatflowMPM(**kw, precision="mixed") for _, kw in scenes]
    with HeatflowBatch(sims) as batch:
        for step in range(3):
            batch.substep(dt)
            for i, a in alone.items():
                a.substep(dt)
                for f in ("T", "U", "V", "W", "P"):
                    assert np.array_equal(a._scene.get_field(f), sims[i]._scene.get_field(f)), (step, i, f)
                assert a.stats()["p_outer"] == sims[i].stats()["p_outer"], (step, i)
    for s in list(alone.values()) + sims:
        s.close()
