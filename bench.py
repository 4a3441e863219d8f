"""Benchmark of the air/heat substep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: the default coupled scene at n=96 (0.48 m box, h = 5 mm, a static ellipsoid of ~26.8k
particles at 20 C, 400 W heater capped at 120 C with a 2 m/s fan, pressure and heat PCG to a
relative residual of 1e-6, dt = 2 ms).  One step is one substep (HeatflowMPM.substep without the
MPM constitutive solve).  W substeps are run untimed, then K substeps are timed between device
synchronisations; value = n^3 * K / seconds (cell-substeps/s).

Beside the throughput the JSON line carries:
  roofline     the pressure-PCG kernel that takes most time, from HIP events around each of its
               launches over a second timed region (same K): algorithmic bytes of one PCG iteration
               of that kernel x real iterations / total kernel time, against 8 TB/s.
  cpu_baseline the fp32 CPU oracle (C restatement of the reference's substep, one thread) on a
               bounded sample of the same workload: it starts from the GPU state after warmup and
               runs substeps for about 15 s.
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

from heatflow_mpm import HeatflowMPM, coupled_scene  # noqa: E402
from heatflow_mpm import _abi  # noqa: E402

METRIC = "cell-substeps/s + HBM TB/s at 96^3; served p50/p99 vs plain at equal tolerance"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per cell of one PCG iteration, by kernel (fp32 words x 4):
#   dir:   reads z, p_old, alpha_s; writes p, Ap                     -> 5 words
#   upd:   reads x, p, r, Ap, D; writes x, r, z                        -> 8 words
#   fused: reads x, r, p_old, D, alpha_s; writes x, r, p              -> 8 words (whole iteration)
KERNEL_WORDS = {_abi.KT_PCG_DIR: 5, _abi.KT_PCG_UPD: 8, _abi.KT_PCG_FUSED: 8}
KERNEL_NAMES = {_abi.KT_PCG_DIR: "k_pcg_dir_tile3", _abi.KT_PCG_UPD: "k_pcg_upd<1>",
                _abi.KT_PCG_FUSED: "k_pcg_fused"}
TRAFFIC_FILE = os.path.join(REPO, "profiles", "pcg_pressure_n96_traffic.json")


def measured_traffic(kernel: str, n: int):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    (calibrated against a plain copy), when they were taken on this workload; else None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("n") != n:
        return None
    for name, rec in t.get("kernels", {}).items():
        if name.startswith(kernel):
            return rec["traffic_bytes"]
    return None


def substep_bytes(N: int, P: int, I_p: float, I_h: float, fused: bool) -> float:
    """Algorithmic HBM bytes of one substep of this implementation (fp32 words x 4):
    P2G heat-sum gather 2 (static particles), heat prep 6.25, T advection + kappa + heat rhs/diag
    10.25, heat PCG init 6.25, heat PCG 13.25 per iteration, air face pass 14, divergence +
    diagonal 9, rhs mean 1, pressure PCG init 7 (+5 for the fused form's first scalars), pressure
    PCG 13 (two-pass) or 8 (fused) per iteration, projection 7 words per cell; 8 words per
    particle for G2P."""
    p_init, p_iter = (12, 8) if fused else (7, 13)
    words = N * (2 + 6.25 + 10.25 + 6.25 + 13.25 * I_h + 14 + 9 + 1 + p_init + p_iter * I_p + 7) + 8 * P
    return 4.0 * words


def cpu_baseline(kw, dt, state, budget_s: float = 15.0):
    """fp32 CPU oracle from the GPU state after warmup; substeps until ~budget_s elapsed."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    orc = O.OracleSim(**kw)
    for name in ("T", "U", "V", "W", "P"):
        orc.set_field(name, state[name])
    if orc.n_particles:
        orc.set_particle_temperatures(state["temp"])
    n = orc.n
    done, t0 = 0, time.perf_counter()
    while True:
        orc.substep(dt)
        done += 1
        el = time.perf_counter() - t0
        if el >= budget_s or done >= 50:
            break
    return dict(value=n ** 3 * done / el, unit="cell-substeps/s", cores=1, kind="port",
                sample=f"n={n} coupled scene, {done} substeps starting from the GPU state after warmup, "
                       f"fp32 C oracle on one host thread, {el:.1f} s")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=96)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cg", choices=HeatflowMPM.CG_FORMS, default="warp",
                    help="pressure PCG form: warp's two-pass iteration, or one fused launch per iteration")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the substep runs on one device (scene-level sharding across GPUs is "
              "independent scenes, not one scene split); run with --gpus 1", file=sys.stderr)
        return 2

    kw, dt, _ = coupled_scene(a.n)
    sim = HeatflowMPM(**kw, device="cuda:0", cg=a.cg)
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
    hbm_tbs = substep_bytes(N, P, I_p, I_h, a.cg == "fused") * a.steps / wall / 1e12

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
    bytes_per_launch = KERNEL_WORDS[dom] * 4.0 * N
    achieved = bytes_per_launch * iters / (total_ms * 1e-3) / 1e9
    traffic = measured_traffic(KERNEL_NAMES[dom], n)
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4),
                    traffic=None if traffic is None else round(traffic), kernel=KERNEL_NAMES[dom],
                    avg_launch_us=round(total_ms * 1e3 / max(launches, 1), 3), launches=launches,
                    iterations=iters, bytes_per_launch=bytes_per_launch)

    cpu = None if a.no_cpu else cpu_baseline(kw, dt, state)

    out = dict(metric=METRIC, value=value, unit="cell-substeps/s", n_gpus=1, steps=a.steps, warmup=a.warmup,
               ms_per_step=wall * 1e3 / a.steps, higher_is_better=True, scaling="weak", vs_baseline=None,
               dtype="f32", data="synthetic",
               config=dict(workload=f"coupled scene n={n} (h=5 mm), static ellipsoid P={P}, heater 400 W cap 120 C fan "
                                    f"2 m/s, PCG tol 1e-6, dt 2 ms", n=n, particles=P, dt=dt, cg_tol=1e-6,
                           pressure_cg=a.cg,
                           pressure_iters_per_substep=round(I_p, 2), heat_iters_per_substep=round(I_h, 2),
                           parallelism="1 GPU"),
               hbm_tb_s=round(hbm_tbs, 3), roofline=roofline, cpu_baseline=cpu)
    print(json.dumps(out), flush=True)
    sim.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
