"""Benchmark of the air/heat substep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cg warp|fused] [--precision fp32|mixed]
    python bench.py --workload batched [--scenes 64] [--steps 200]
    python bench.py --workload served [--rate 20] [--requests 24] [--precision mixed|fp32] [--fp64-n 64]
    python bench.py --workload mixed [--n 192] [--steps 100] [--fp64-n 48]

Workload: the default coupled scene at n=96 (0.48 m box, h = 5 mm, a static ellipsoid of ~26.8k
particles at 20 C, 400 W heater capped at 120 C with a 2 m/s fan, pressure and heat PCG to a
relative residual of 1e-6, dt = 2 ms).  One step is one substep (HeatflowMPM.substep without the
MPM constitutive solve).  W substeps are run untimed, then K substeps are timed between device
synchronisations; value = n^3 * K / seconds (cell-substeps/s).

Beside the throughput the JSON line carries:
  roofline     the pressure-PCG kernel that takes most time, from HIP events around each of its
               launches over a second timed region (same K): algorithmic bytes of one PCG iteration
               of that kernel x real iterations / total kernel time, against 8 TB/s.
  cpu_baseline the fp32 CPU oracle (C restatement of the reference's substep) on a bounded sample
               of the same workload: it starts from the GPU state after warmup and runs substeps
               for about 15 s on one host thread (value, cores = 1) and about 8 s on all OpenMP
               threads (parallel).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "heatflow-mpm_amd"))

import numpy as np  # noqa: E402

from heatflow_mpm import HeatflowBatch, HeatflowMPM, batched_scenes, coupled_scene  # noqa: E402
from heatflow_mpm.serve import Server, poisson_requests, summarize  # noqa: E402
from heatflow_mpm import _abi  # noqa: E402

METRIC = "cell-substeps/s + HBM TB/s at 96^3; served p50/p99 vs plain at equal tolerance"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per cell of one PCG iteration, by kernel (fp32 words x 4):
#   dir:   reads z, p_old, alpha_s; writes p, Ap                     -> 5 words
#   upd:   reads x, p, r, Ap, D; writes x, r, z                        -> 8 words
#   fused: reads x, r, p_old, s_old, w, D, alpha_s; writes x, r, p, s, w -> 12 words (whole iteration)
#   persistent: the fused iteration's algorithmic bytes, per iteration of its one launch per solve (it
#   keeps the vectors on chip and moves only tile faces and partial sums through memory)
KERNEL_WORDS = {_abi.KT_PCG_DIR: 5, _abi.KT_PCG_UPD: 8, _abi.KT_PCG_FUSED: 12, _abi.KT_PCG_PERSIST: 12}
# Mixed precision, inner solve (bytes per cell): pass A reads z, p_old (FP16) and alpha_s, writes
# p, Ap (FP16) -> 12 B; pass B reads x (fp32), p, r, Ap (FP16), D (fp32), writes x, r, z -> 22 B
KERNEL_BYTES_MIXED = {_abi.KT_PCG_DIR: 12, _abi.KT_PCG_UPD: 22}
KERNEL_NAMES = {_abi.KT_PCG_DIR: "k_pcg_dir_tile3", _abi.KT_PCG_UPD: "k_pcg_upd<1>",
                _abi.KT_PCG_FUSED: "k_pcg_fused", _abi.KT_PCG_PERSIST: "k_pcg_persist"}
TRAFFIC_FILE = os.path.join(REPO, "profiles", "pcg_pressure_n96_traffic.json")


def measured_traffic(kernel: str, n: int):
    """Bytes per launch of `kernel` that the memory side (TCC) moved, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE runs of this workload (scaled by what the counters read for a plain copy
    of known size), with where they come from; None if not measured for this kernel and size.  A
    stored value, not measured in this run; at 96^3 the PCG working set fits the 256 MB Infinity
    Cache, so part of these bytes are served by it rather than by HBM."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("n") != n:
        return None
    for name, rec in t.get("kernels", {}).items():
        if name.startswith(kernel):
            return dict(traffic_bytes=rec["traffic_bytes"],
                        source=f"{os.path.relpath(TRAFFIC_FILE, REPO)} (stored rocprofv3 FETCH_SIZE/WRITE_SIZE, "
                               f"{rec.get('measured', t.get('measured', 'earlier run'))}"
                               f"{', per iteration' if rec.get('per') == 'iteration' else ''}; at this size part of "
                               f"it is Infinity Cache traffic, not HBM)")
    return None


def substep_bytes(N: int, P: int, I_p: float, I_h: float, fused: bool) -> float:
    """Algorithmic HBM bytes of one substep of this implementation (fp32 words x 4):
    P2G heat-sum gather 2 (static particles), heat prep 6.25, T advection + kappa + heat rhs/diag
    10.25, heat PCG init 6.25, heat PCG 13.25 per iteration, air face pass 14, divergence +
    diagonal 9, rhs mean 1, pressure PCG init 7 (+3 for the one-launch form's w = A D r), pressure
    PCG 13 (two-pass) or 12 (one launch) per iteration, projection 7 words per cell; 8 words per
    particle for G2P."""
    p_init, p_iter = (10, 12) if fused else (7, 13)
    words = N * (2 + 6.25 + 10.25 + 6.25 + 13.25 * I_h + 14 + 9 + 1 + p_init + p_iter * I_p + 7) + 8 * P
    return 4.0 * words


def contract_bytes(N: int, P: int, I_p: float, I_h: float) -> float:
    """The path's algorithmic bytes per substep as the scope table fixes them (independent of this
    implementation's kernel split): B = 4 [N (48.5 + 12 I_p + 12.5 I_h) + 17 P], with I_p and I_h the
    exact pressure and heat PCG iterations."""
    return 4.0 * (N * (48.5 + 12.0 * I_p + 12.5 * I_h) + 17.0 * P)


def cpu_baseline(kw, dt, state, budget_s: float = 15.0):
    """fp32 CPU oracle from the GPU state after warmup, substeps until ~budget_s elapsed: once on
    one host thread (the reported value) and once on every thread OpenMP is given (the oracle's
    per-cell loops in parallel; sums stay serial, results are the same bits)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    lib = O.lib("f32")
    threads_all = int(lib.orc_max_threads())

    def run(threads, budget):
        lib.orc_set_threads(threads)
        orc = O.OracleSim(**kw)
        for name in ("T", "U", "V", "W", "P"):
            orc.set_field(name, state[name])
        if orc.n_particles:
            orc.set_particle_temperatures(state["temp"])
        done, t0 = 0, time.perf_counter()
        while True:
            orc.substep(dt)
            done += 1
            el = time.perf_counter() - t0
            if el >= budget or done >= 50:
                break
        return orc.n, done, el

    n, done, el = run(1, budget_s)
    _, done_p, el_p = run(threads_all, budget_s / 2)
    lib.orc_set_threads(threads_all)
    return dict(value=n ** 3 * done / el, unit="cell-substeps/s", cores=1, kind="port",
                sample=f"n={n} coupled scene, {done} substeps starting from the GPU state after warmup, "
                       f"fp32 C oracle on one host thread, {el:.1f} s",
                parallel=dict(value=n ** 3 * done_p / el_p, cores=threads_all,
                              sample=f"same, {done_p} substeps on {threads_all} OpenMP threads, {el_p:.1f} s"))


def bench_batched(a) -> dict:
    """Uneven independent scenes (n in {32..128}, random heaters, fans and particle blobs), all
    advanced K substeps from rest: once one scene at a time (the plain form), once as one batch
    (every stage one launch over all scenes).  Same work both ways; the final states are compared
    scene by scene (they must be bit-identical)."""
    scenes, dt, K0 = batched_scenes(a.scenes, seed=a.seed)
    K = a.steps if a.steps_given else K0
    cells = sum(kw["grid"].n ** 3 for _, kw in scenes)
    # load the code objects of the solo and the batched kernels outside the timed regions
    warm = [HeatflowMPM(**coupled_scene(32)[0], cg=a.cg) for _ in range(2)]
    warm[0].run(dt, 2)
    with HeatflowBatch(warm) as wb:
        wb.run(dt, 2)
    for w in warm:
        w.close()
    prec = a.precision or "fp32"
    lay = lambda ctx: None if a.layout == "rule" else ("auto-batch" if ctx == "batch" else "auto")  # noqa: E731
    plain = [HeatflowMPM(**kw, cg=a.cg, precision=prec, layout=lay("solo")) for _, kw in scenes]
    t_plain, p_iters_plain = 0.0, 0
    for s in plain:
        s.synchronize()
        t0 = time.perf_counter()
        s.run(dt, K)
        s.synchronize()
        t_plain += time.perf_counter() - t0
        p_iters_plain += s.stats()["p_iters_total"]
    sims = [HeatflowMPM(**kw, cg=a.cg, precision=prec, layout=lay("batch")) for _, kw in scenes]
    batch = HeatflowBatch(sims)
    batch.synchronize()
    t0 = time.perf_counter()
    batch.run(dt, K)
    batch.synchronize()
    t_batch = time.perf_counter() - t0
    # a scene gives the same bits alone and in the batch for the same layout (a different pass-A
    # tiling groups the partial sums differently): compare the scenes whose tuned layouts agree
    same, compared = True, 0
    for s, p in zip(sims, plain):
        if s.layout != p.layout:
            continue
        compared += 1
        for f in ("T", "U", "V", "W", "P"):
            if not np.array_equal(s._scene.get_field(f), p._scene.get_field(f)):
                same = False
        if not np.array_equal(s.temperatures(), p.temperatures()):
            same = False
    value, plain_value = cells * K / t_batch, cells * K / t_plain
    ns = [kw["grid"].n for _, kw in scenes]
    out = dict(metric="aggregate cell-substeps/s, batched uneven scenes", value=value, unit="cell-substeps/s",
               n_gpus=1, steps=K, warmup=0, ms_per_step=t_batch * 1e3 / K, higher_is_better=True, scaling="weak",
               vs_baseline=None, dtype="f32", data="synthetic",
               config=dict(workload=f"{len(scenes)} independent scenes, n in {{32,48,64,96,128}} "
                                    f"(counts {[ns.count(m) for m in (32, 48, 64, 96, 128)]}), heater 100-800 W, "
                                    f"fan 0-3 m/s, 2k-50k particles, PCG tol 1e-6, dt 2 ms, {K} substeps from rest",
                           scenes=len(scenes), cells=cells, seed=a.seed, pressure_cg=a.cg, precision=prec,
                           layout=a.layout,
                           parallelism="1 GPU"),
               plain=dict(value=plain_value, seconds=round(t_plain, 3), pressure_iters_total=int(p_iters_plain)),
               speedup_vs_plain=round(value / plain_value, 3), batched_seconds=round(t_batch, 3),
               identical_to_plain=same, scenes_compared=compared)
    batch.close()
    for s in sims + plain:
        s.close()
    return out


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))


def bench_served(a) -> dict:
    """A seeded Poisson stream of rollout requests (n from {48, 64, 96, 128, 192}, K uniform in
    [kmin, kmax], heaters / fans / particle blobs as the batched scenes), served twice on the same
    arrival times: with continuous batching (the admitted scenes of a size class advance in one batched
    substep, the size classes (--lanes) side by side on their own streams; admission in arrival order by
    device footprint), the multigrid-preconditioned pressure PCG and the mixed-precision form (FP16
    refinement vectors, fp32 state) by default, and in the plain form (fp32,
    the reference's Jacobi PCG, one request at a time, in arrival order).
    Latency = completion - arrival.  Each served result is compared with
    the plain result of the same request (relative L2 of T and u): both solve to the same relative
    residual, so they agree to the solve tolerance, not in bits."""
    reqs = poisson_requests(a.rate, a.requests, seed=a.seed, k_range=(a.kmin, a.kmax), cg_max_iter=a.cg_max_iter)
    scg = a.cg or "mg"
    warm = [HeatflowMPM(**coupled_scene(32)[0], cg=scg) for _ in range(2)]
    warm[0].run(2e-3, 2)
    with HeatflowBatch(warm) as wb:
        wb.run(2e-3, 2)
    for w in warm:
        w.close()
    from heatflow_mpm.serve import GpuBackend
    budget = int(a.hbm_gib * (1 << 30))
    prec = a.precision or "mixed"   # the served configuration names the mixed-precision form
    served = Server(GpuBackend(cg=scg, precision=prec, layout=None if a.layout == "rule" else "auto-batch"),
                    hbm_budget=budget, max_active=a.slots, keep_results=True, lanes=_lanes(a))
    t0 = time.perf_counter()
    res = served.run(reqs)
    span = time.perf_counter() - t0
    s_b = summarize(res, span)
    plain = Server(GpuBackend(cg=a.plain_cg, layout=None if a.layout == "rule" else "auto"), hbm_budget=budget,
                   max_active=1, keep_results=True)
    t0 = time.perf_counter()
    res_p = plain.run(reqs)
    span_p = time.perf_counter() - t0
    s_p = summarize(res_p, span_p)
    ns = [r.kwargs["grid"].n for r in reqs]
    err_T = err_u = 0.0
    per = []
    for rb, rp in zip(res, res_p):
        eT = rel_l2(rb.fields["T"], rp.fields["T"])
        ub = np.concatenate([rb.fields[f].ravel() for f in ("U", "V", "W")])
        up = np.concatenate([rp.fields[f].ravel() for f in ("U", "V", "W")])
        eu = rel_l2(ub, up)
        err_T, err_u = max(err_T, eT), max(err_u, eu)
        per.append(dict(rid=rb.rid, n=rb.n, substeps=rb.substeps, T=eT, u=eu, p_converged=rb.stats.get("p_converged")))
    worst = sorted(per, key=lambda d: -d["u"])[:4]
    # error of every served result against the fp64 CPU oracle, for the requests whose oracle run fits (n <= fp64-n)
    vs_fp64 = []
    if a.fp64_n:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O
        for rb, q in zip(res, reqs):
            if rb.n > a.fp64_n:
                continue
            orc = O.OracleSim(**q.kwargs, precision="f64")
            orc.run(q.dt, q.substeps)
            uo = np.concatenate([orc.field(f).ravel() for f in ("U", "V", "W")])
            ub = np.concatenate([rb.fields[f].ravel() for f in ("U", "V", "W")])
            d64, _ = orc.diagnostics()
            vs_fp64.append(dict(rid=rb.rid, n=rb.n, substeps=rb.substeps, T=rel_l2(rb.fields["T"], orc.field("T")),
                                u=rel_l2(ub, uo), divergence_max=rb.stats.get("divergence_max"),
                                divergence_max_fp64=d64))
            del orc
    return dict(metric="served p50/p99 latency and aggregate cell-substeps/s, continuous batching",
                value=s_b["cell_substeps_per_s"], unit="cell-substeps/s", n_gpus=1, steps=None, warmup=0,
                ms_per_step=None, higher_is_better=True, scaling="weak", vs_baseline=None,
                dtype="f32" if prec == "fp32" else "f32 state, f16 PCG vectors",
                data="synthetic",
                config=dict(workload=f"{len(reqs)} Poisson requests at {a.rate}/s, n in {{48,64,96,128,192}} "
                                     f"(counts {[ns.count(m) for m in (48, 64, 96, 128, 192)]}), K uniform "
                                     f"{a.kmin}-{a.kmax}, heater 100-800 W, fan 0-3 m/s, 2k-50k particles, "
                                     f"PCG tol 1e-6 (pressure cap {a.cg_max_iter}), dt 2 ms", rate_per_s=a.rate,
                            requests=len(reqs), seed=a.seed, cg_max_iter=a.cg_max_iter,
                            hbm_budget_gib=a.hbm_gib, slots=a.slots, lanes=a.lanes, pressure_cg=scg,
                            plain_cg=a.plain_cg, precision=prec,
                            layout=a.layout,
                            parallelism="1 GPU"),
                max_rel_l2_vs_plain=dict(T=err_T, u=err_u), worst_requests_vs_plain=worst,
                vs_fp64_oracle=vs_fp64 or None,
                served=s_b, plain=s_p, speedup_vs_plain=round(s_b["cell_substeps_per_s"] / s_p["cell_substeps_per_s"], 3),
                p99_ratio_plain_over_served=round(s_p["p99_s"] / s_b["p99_s"], 3),
                max_batch=max(e[3] for e in served.log))


def _lanes(a):
    """Size-class lanes of the served form (MAXN:SLOTS each); "none" serves one batch of all requests."""
    if not a.lanes or a.lanes == ["none"]:
        return None
    return [tuple(int(x) for x in lane.split(":")) for lane in a.lanes]


def bench_served_sweep(a) -> dict:
    """The served stream at several arrival rates (the same seeded requests: the exponential gaps
    scale with 1/rate, the scene draws do not change), served and plain at each rate."""
    from heatflow_mpm.serve import GpuBackend
    scg = a.cg or "mg"
    prec = a.precision or "mixed"   # the served configuration names the mixed-precision form
    budget = int(a.hbm_gib * (1 << 30))
    warm = [HeatflowMPM(**coupled_scene(32)[0], cg=scg) for _ in range(2)]
    with HeatflowBatch(warm) as wb:
        wb.run(2e-3, 2)
    for w in warm:
        w.close()
    rows = []
    for rate in a.rates:
        reqs = poisson_requests(rate, a.requests, seed=a.seed, k_range=(a.kmin, a.kmax), cg_max_iter=a.cg_max_iter)
        served = Server(GpuBackend(cg=scg, precision=prec, layout=None if a.layout == "rule" else "auto-batch"),
                        hbm_budget=budget, max_active=a.slots, lanes=_lanes(a))
        t0 = time.perf_counter()
        s_b = summarize(served.run(reqs), time.perf_counter() - t0)
        plain = Server(GpuBackend(cg=a.plain_cg, layout=None if a.layout == "rule" else "auto"), hbm_budget=budget,
                       max_active=1)
        t0 = time.perf_counter()
        s_p = summarize(plain.run(reqs), time.perf_counter() - t0)
        rows.append(dict(rate_per_s=rate, served=s_b, plain=s_p,
                         speedup=round(s_b["cell_substeps_per_s"] / s_p["cell_substeps_per_s"], 3),
                         p99_ratio=round(s_p["p99_s"] / s_b["p99_s"], 3)))
        print(f"rate {rate}/s: served {s_b['cell_substeps_per_s']:.3e} cell-substeps/s p50 {s_b['p50_s']:.2f} s "
              f"p99 {s_b['p99_s']:.2f} s; plain {s_p['cell_substeps_per_s']:.3e} p50 {s_p['p50_s']:.2f} "
              f"p99 {s_p['p99_s']:.2f}", file=sys.stderr, flush=True)
    return dict(metric="served p50/p99 latency and aggregate cell-substeps/s against arrival rate",
                value=min(r["speedup"] for r in rows), unit="x plain aggregate cell-substeps/s (lowest over rates)",
                n_gpus=1, steps=None, warmup=0, ms_per_step=None, higher_is_better=True, scaling="weak",
                vs_baseline=None, dtype="f32" if prec == "fp32" else "f32 state, f16 PCG vectors",
                data="synthetic",
                config=dict(workload=f"{a.requests} Poisson requests per rate, n in {{48,64,96,128,192}}, K uniform "
                                     f"{a.kmin}-{a.kmax}, PCG tol 1e-6 (pressure cap {a.cg_max_iter}), dt 2 ms",
                            rates=a.rates, seed=a.seed, pressure_cg=scg, plain_cg=a.plain_cg, precision=prec,
                            hbm_budget_gib=a.hbm_gib, slots=a.slots, lanes=a.lanes, parallelism="1 GPU"),
                sweep=rows)


def bench_mixed(a) -> dict:
    """The coupled scene at n=192 (h = 2.5 mm) with the ellipsoid seeded to about 50k particles,
    K substeps from rest, once in fp32 and once in the mixed form (iterative refinement with FP16
    PCG vectors) on the same GPU.  Reports both rates, the pressure-PCG
    kernel times, and the relative L2 of the mixed T and u against fp32; with --fp64-n m also the
    error of both forms against the fp64 CPU oracle after K substeps of the same scene at n = m."""
    import dataclasses
    n = a.n if a.n_given else 192
    K = a.steps if a.steps_given else 100
    kw, dt, _ = coupled_scene(n, particles_target=50000)
    # Jacobi-PCG at n = 192 needs more than the reference's default 400 iterations to reach 1e-6
    kw["grid"] = dataclasses.replace(kw["grid"], cg_max_iter=a.cg_max_iter)
    out_forms = {}
    fields = {}
    for form in ("fp32", "mixed"):
        sim = HeatflowMPM(**kw, precision=form)
        sim.synchronize()
        t0 = time.perf_counter()
        sim.run(dt, K)
        sim.synchronize()
        wall = time.perf_counter() - t0
        st = sim.stats()
        dm, sm = sim._scene.diagnostics()
        fields[form] = {f: sim._scene.get_field(f) for f in ("T", "U", "V", "W", "P")}
        # kernel times from HIP events over a few more substeps (events are kept out of the timed run)
        sim._scene.set_kernel_timing(True)
        sim.run(dt, min(K, 10))
        sim.synchronize()
        kt = {c: sim._scene.kernel_time(c) for c in (_abi.KT_PCG_DIR, _abi.KT_PCG_UPD)}
        sim._scene.set_kernel_timing(False)
        out_forms[form] = dict(value=n ** 3 * K / wall, seconds=round(wall, 3), ms_per_step=wall * 1e3 / K,
                               pressure_iters_per_substep=st["p_iters_total"] / K,
                               heat_iters_per_substep=st["h_iters_total"] / K,
                               pass_a_us=round(kt[_abi.KT_PCG_DIR][0] * 1e3 / max(kt[_abi.KT_PCG_DIR][1], 1), 2),
                               pass_b_us=round(kt[_abi.KT_PCG_UPD][0] * 1e3 / max(kt[_abi.KT_PCG_UPD][1], 1), 2),
                               divergence_max=dm, speed_max=sm, p_converged=st["p_converged"],
                               p_outer_last=st["p_outer"], h_outer_last=st["h_outer"], particles=sim.n)
        sim.close()
    fm, f32 = fields["mixed"], fields["fp32"]
    u_m = np.concatenate([fm[f].ravel() for f in "UVW"])
    u_32 = np.concatenate([f32[f].ravel() for f in "UVW"])
    vs_fp32 = dict(T=rel_l2(fm["T"], f32["T"]), u=rel_l2(u_m, u_32),
                   p_centered=rel_l2(fm["P"] - fm["P"].mean(), f32["P"] - f32["P"].mean()))
    vs_fp64 = None
    if a.fp64_n:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle as O
        kw_s, _, _ = coupled_scene(a.fp64_n, particles_target=int(50000 * (a.fp64_n / 192) ** 3) or None)
        kw_s["grid"] = dataclasses.replace(kw_s["grid"], cg_max_iter=a.cg_max_iter)
        ref = O.OracleSim(**kw_s, precision="f64")
        ref.run(dt, K)
        uref = np.concatenate([ref.field(f).ravel() for f in "UVW"])
        vs_fp64 = dict(n=a.fp64_n, substeps=K)
        for form in ("fp32", "mixed"):
            sim = HeatflowMPM(**kw_s, precision=form)
            sim.run(dt, K)
            sim.synchronize()
            ug = np.concatenate([sim._scene.get_field(f).ravel() for f in "UVW"])
            dm, _ = sim._scene.diagnostics()
            vs_fp64[form] = dict(T=rel_l2(sim._scene.get_field("T"), ref.field("T")), u=rel_l2(ug, uref),
                                 divergence_max=dm)
            sim.close()
        vs_fp64["fp64_divergence_max"] = ref.diagnostics()[0]
    mixed = out_forms["mixed"]
    return dict(metric="cell-substeps/s, mixed precision (iterative refinement, FP16 PCG vectors)",
                value=mixed["value"], unit="cell-substeps/s", n_gpus=1, steps=K, warmup=0,
                ms_per_step=mixed["ms_per_step"], higher_is_better=True, scaling="weak", vs_baseline=None,
                dtype="f32 state, f16 PCG vectors", data="synthetic",
                config=dict(workload=f"coupled scene n={n} (h={0.48 / n * 1e3:.2f} mm), static ellipsoid "
                                     f"P={mixed['particles']}, heater 400 W cap 120 C fan 2 m/s, PCG tol 1e-6 "
                                     f"(pressure cap {a.cg_max_iter} iterations), dt 2 ms, {K} substeps from rest",
                            n=n, cg_max_iter=a.cg_max_iter, parallelism="1 GPU"),
                fp32=out_forms["fp32"], mixed=mixed,
                speedup_vs_fp32=round(mixed["value"] / out_forms["fp32"]["value"], 3),
                rel_l2_mixed_vs_fp32=vs_fp32, error_vs_fp64=vs_fp64)


def pcg_kernel_fractions(n: int, warmup: int, steps: int) -> dict:
    """The plain form's pressure-PCG passes on the coupled scene at a larger grid, event-timed like
    the main roofline: per kernel the average in-dispatch time, the algorithmic GB/s and the
    fraction of HBM peak (at 96^3 the passes are latency-bound; here they are not)."""
    kw, dt, _ = coupled_scene(n, particles_target=50000)
    import dataclasses
    kw["grid"] = dataclasses.replace(kw["grid"], cg_max_iter=2000)
    sim = HeatflowMPM(**kw, device="cuda:0", layout="auto")
    sim.run(dt, warmup)
    sim.synchronize()
    sim._scene.set_kernel_timing(True)
    sim._scene.kernel_time_reset()
    st0 = sim.stats()
    sim.run(dt, steps)
    sim.synchronize()
    iters = sim.stats()["p_iters_total"] - st0["p_iters_total"]
    out = {}
    for c in (_abi.KT_PCG_DIR, _abi.KT_PCG_UPD):
        total_ms, launches = sim._scene.kernel_time(c)
        if launches:
            gbs = KERNEL_WORDS[c] * 4.0 * n ** 3 * iters / (total_ms * 1e-3) / 1e9
            out[KERNEL_NAMES[c]] = dict(avg_launch_us=round(total_ms * 1e3 / launches, 2), achieved=round(gbs, 1),
                                        unit="GB/s", frac=round(gbs / HBM_PEAK_GBS, 3))
    sim._scene.set_kernel_timing(False)
    sim.close()
    return dict(n=n, substeps=steps, pressure_iters=iters, kernels=out)


def bench_mg_form(a, kw, dt, end) -> dict:
    """The same scene and substeps with the multigrid-preconditioned pressure PCG (fp32): warmup,
    the timed substeps, then as many untimed ones again, so that its T and u stand beside the plain
    run's after the same substeps (the plain run takes a second, kernel-timed pass)."""
    sim = HeatflowMPM(**kw, device="cuda:0", cg="mg", layout=None if a.layout == "rule" else "auto")
    N = sim.grid.n ** 3
    if a.warmup:
        sim.run(dt, a.warmup)
    st0 = sim.stats()
    sim.synchronize()
    t0 = time.perf_counter()
    sim.run(dt, a.steps)
    sim.synchronize()
    wall = time.perf_counter() - t0
    st1 = sim.stats()
    sim.run(dt, a.steps)
    sim.synchronize()
    u = np.concatenate([sim._scene.get_field(f).ravel() for f in ("U", "V", "W")])
    u_plain = np.concatenate([end[f].ravel() for f in ("U", "V", "W")])
    out = dict(value=N * a.steps / wall, ms_per_step=wall * 1e3 / a.steps,
               pressure_iters_per_substep=round((st1["p_iters_total"] - st0["p_iters_total"]) / a.steps, 2),
               rel_l2_vs_plain=dict(T=rel_l2(sim._scene.get_field("T"), end["T"]), u=rel_l2(u, u_plain)),
               substeps_compared=a.warmup + 2 * a.steps)
    sim.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("coupled", "batched", "served", "mixed"), default="coupled",
                    help="coupled: one n=96 coupled scene (default); batched: uneven scenes in one batch vs one by "
                         "one; served: a Poisson request stream with continuous batching vs one at a time; "
                         "mixed: the n=192 scene in fp32 and in mixed precision")
    ap.add_argument("--precision", choices=tuple(HeatflowMPM.PRECISIONS), default=None,
                    help="coupled/batched: precision of the substep (fp32 default); served: precision of the "
                         "batched server (mixed default; the plain form is fp32)")
    ap.add_argument("--layout", choices=("rule", "auto"), default="auto",
                    help="pressure PCG layout: chosen by grid size, or the autotuned table (heatflow_mpm.autotune)")
    ap.add_argument("--fp64-n", type=int, default=0,
                    help="mixed: grid size of the fp64-oracle error check; served: check every result of at most this "
                         "grid size against the fp64 oracle (0: off)")
    ap.add_argument("--cg-tol", type=float, default=None,
                    help="coupled: pressure PCG tolerance (default 1e-6 of the configuration; the reference's own "
                         "default is 1e-4)")
    ap.add_argument("--cg-max-iter", type=int, default=2000,
                    help="mixed and served: pressure PCG iteration cap (n=192 needs more than the default 400)")
    ap.add_argument("--rate", type=float, default=20.0, help="served: arrivals per second")
    ap.add_argument("--rates", type=float, nargs="+", default=None,
                    help="served: sweep these arrival rates (one JSON line with a row per rate)")
    ap.add_argument("--requests", type=int, default=24, help="served: number of requests")
    ap.add_argument("--kmin", type=int, default=20, help="served: fewest substeps of a request")
    ap.add_argument("--kmax", type=int, default=100, help="served: most substeps of a request")
    ap.add_argument("--hbm-gib", type=float, default=64.0, help="served: device memory budget of the scenes")
    ap.add_argument("--slots", type=int, default=64, help="served: most scenes in one batch")
    ap.add_argument("--lanes", nargs="+", default=["64:64", "128:8", "1024:1"], metavar="MAXN:SLOTS",
                    help="served: size-class lanes served side by side, each its own continuous batch on its own "
                         "stream (default: n <= 64 in batches of up to 64, n <= 128 of up to 8, larger ones one at a "
                         "time); 'none' for one batch of every request")
    ap.add_argument("--scenes", type=int, default=64, help="batched: number of scenes")
    ap.add_argument("--seed", type=int, default=0, help="batched: scene draw seed")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=None, help="grid size (coupled: 96, mixed: 192)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-large", action="store_true",
                    help="coupled: skip the event-timed PCG passes of the same scene at 192^3")
    ap.add_argument("--no-mg", action="store_true",
                    help="coupled: skip the multigrid-preconditioned form timed beside the plain one")
    ap.add_argument("--cg", choices=HeatflowMPM.CG_FORMS, default=None,
                    help="pressure PCG form: warp's two-pass Jacobi iteration (default), one fused launch per "
                         "iteration, or the multigrid-preconditioned iteration (mg; the served form's default)")
    ap.add_argument("--advection", choices=HeatflowMPM.ADVECTION, default="semi-lagrangian",
                    help="coupled: the reference's semi-Lagrangian step or limited MacCormack")
    ap.add_argument("--plain-cg", choices=HeatflowMPM.CG_FORMS, default="warp",
                    help="served: pressure PCG form of the plain baseline (the reference's Jacobi PCG)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the substep runs on one device (scene-level sharding across GPUs is "
              "independent scenes, not one scene split); run with --gpus 1", file=sys.stderr)
        return 2
    if a.workload != "served" and a.cg is None:
        a.cg = "warp"
    a.steps_given = a.steps is not None
    a.n_given = a.n is not None
    if a.n is None:
        a.n = 96
    if a.workload == "mixed":
        print(json.dumps(bench_mixed(a)), flush=True)
        return 0
    if a.workload == "batched":
        print(json.dumps(bench_batched(a)), flush=True)
        return 0
    if a.workload == "served":
        print(json.dumps(bench_served_sweep(a) if a.rates else bench_served(a)), flush=True)
        return 0
    if a.steps is None:
        a.steps = 500

    prec = a.precision or "fp32"
    kw, dt, _ = coupled_scene(a.n)
    if a.cg_tol is not None:
        import dataclasses
        kw["grid"] = dataclasses.replace(kw["grid"], cg_tol=a.cg_tol)
    sim = HeatflowMPM(**kw, device="cuda:0", cg=a.cg, precision=prec, layout=None if a.layout == "rule" else "auto",
                      advection=a.advection)
    n, N, P = sim.grid.n, sim.grid.n ** 3, sim.n
    if a.warmup:
        sim.run(dt, a.warmup)
    sim.synchronize()
    state = {name: sim._scene.get_field(name) for name in ("T", "U", "V", "W", "P")}
    state["temp"] = sim.temperatures()
    st0 = sim.stats()

    # timed region
    sim.synchronize()
    t0 = time.perf_counter()
    sim.run(dt, a.steps)
    sim.synchronize()
    wall = time.perf_counter() - t0
    st1 = sim.stats()
    I_p = (st1["p_iters_total"] - st0["p_iters_total"]) / a.steps
    I_h = (st1["h_iters_total"] - st0["h_iters_total"]) / a.steps
    value = N * a.steps / wall
    hbm_impl = (substep_bytes(N, P, I_p, I_h, sim.layout["cg"] in ("fused", "persistent")) * a.steps / wall / 1e12
                if prec == "fp32" else None)
    B = contract_bytes(N, P, I_p, I_h)
    hbm_tbs = B * a.steps / wall / 1e12 if prec == "fp32" else None
    substep_roofline = (None if prec != "fp32" else
                        dict(bound="hbm", bytes_per_substep=round(B), achieved_tb_s=round(hbm_tbs, 3),
                             peak_tb_s=HBM_PEAK_GBS / 1e3, frac=round(hbm_tbs * 1e3 / HBM_PEAK_GBS, 4),
                             formula="B = 4 [N (48.5 + 12 I_p + 12.5 I_h) + 17 P] per substep, K B / wall time",
                             own_kernel_model_tb_s=None if hbm_impl is None else round(hbm_impl, 3)))

    # second timed region with HIP events around every pressure-PCG launch
    sim._scene.set_kernel_timing(True)
    sim._scene.kernel_time_reset()
    st2 = sim.stats()
    sim.run(dt, a.steps)
    sim.synchronize()
    st3 = sim.stats()
    iters = st3["p_iters_total"] - st2["p_iters_total"]
    kt = {c: sim._scene.kernel_time(c) for c in KERNEL_WORDS}
    sim._scene.set_kernel_timing(False)
    dom = max(kt, key=lambda c: kt[c][0])
    total_ms, launches = kt[dom]
    bytes_per_iteration = (KERNEL_WORDS[dom] * 4.0 if prec == "fp32" else KERNEL_BYTES_MIXED[dom]) * N
    # one launch per iteration, except the persistent form: one launch per solve
    bytes_per_launch = bytes_per_iteration * iters / max(launches, 1)
    achieved = bytes_per_iteration * iters / (total_ms * 1e-3) / 1e9
    tr = measured_traffic(KERNEL_NAMES[dom], n) if prec == "fp32" else None
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4),
                    traffic=None if tr is None else round(tr["traffic_bytes"] * (iters / max(launches, 1)
                                                                                if dom == _abi.KT_PCG_PERSIST
                                                                                else 1.0)),
                    kernel=KERNEL_NAMES[dom],
                    avg_launch_us=round(total_ms * 1e3 / max(launches, 1), 3), launches=launches,
                    iterations=iters, bytes_per_launch=round(bytes_per_launch),
                    bytes_per_iteration=bytes_per_iteration,
                    traffic_source=None if tr is None else tr["source"])
    if dom == _abi.KT_PCG_PERSIST:
        roofline["note"] = ("achieved counts the one-launch iteration's algorithmic bytes (12 words per cell per "
                            "iteration); the persistent kernel keeps the vectors on chip and moves only tile faces "
                            "and partial sums through memory (traffic), so it is bound by instruction issue and the "
                            "grid-wide wait, not by HBM")

    cpu = None if a.no_cpu else cpu_baseline(kw, dt, state)
    end = {name: sim._scene.get_field(name) for name in ("T", "U", "V", "W")}
    mg = (None if a.no_mg or a.cg != "warp" or prec != "fp32" or a.advection != "semi-lagrangian"
          else bench_mg_form(a, kw, dt, end))
    large = (None if a.no_large or a.cg != "warp" or prec != "fp32" or a.n != 96
             else pcg_kernel_fractions(192, 5, 10))

    out = dict(metric=METRIC, value=value, unit="cell-substeps/s", n_gpus=1, steps=a.steps, warmup=a.warmup,
               ms_per_step=wall * 1e3 / a.steps, higher_is_better=True, scaling="weak", vs_baseline=None,
               dtype="f32" if prec == "fp32" else "f32 state, f16 PCG vectors",
               data="synthetic",
               config=dict(workload=f"coupled scene n={n} (h=5 mm), static ellipsoid P={P}, heater 400 W cap 120 C fan "
                                    f"2 m/s, PCG tol {kw['grid'].cg_tol:g}, dt 2 ms", n=n, particles=P, dt=dt,
                           cg_tol=kw["grid"].cg_tol,
                           pressure_cg=sim.layout["cg"], precision=prec, layout=a.layout, advection=a.advection,
                           dir_planes=sim.layout["dir_planes"],
                           pressure_iters_per_substep=round(I_p, 2), heat_iters_per_substep=round(I_h, 2),
                           pressure_iters_since_start=int(st3["p_iters_total"]), parallelism="1 GPU"),
               hbm_tb_s=None if hbm_tbs is None else round(hbm_tbs, 3), substep_roofline=substep_roofline,
               roofline=roofline, cpu_baseline=cpu,
               mg_form=mg, pcg_kernels_192=large)
    print(json.dumps(out), flush=True)
    sim.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
