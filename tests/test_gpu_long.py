"""Long-horizon parity: the GPU substep against the fp32 oracle in lockstep over 500 substeps.

The bar (fp32 T and u within relative L2 1e-3 of the oracle after 500 substeps, CG iteration counts
within one at every substep) is held on the plume scene (n=48, no material; the substep loop of
solver.py:63-88 as tests/test_solver.py:16-18 runs it) and on the coupled scene at n=64 (static
ellipsoid of particles, fan 2 m/s).  The plume is also compared with the fp64 oracle, and the error
of both fp32 runs against it is printed.
"""
import time

import numpy as np
import pytest

import oracle as O
from conftest import rel_l2
from heatflow_mpm import HeatflowMPM, coupled_scene, plume_scene

pytestmark = pytest.mark.gpu

K = 500


def _lockstep(kw, dt, steps, f64=None):
    sim = HeatflowMPM(**kw)
    orc = O.OracleSim(**kw)
    worst_p = worst_h = 0
    t0 = time.time()
    for step in range(steps):
        sim.substep(dt)
        orc.substep(dt)
        if f64 is not None:
            f64.substep(dt)
        st, so = sim.stats(), orc.stats()
        dp = abs(st["p_iters_exact"] - so["p_iters_exact"])
        dh = abs(st["h_iters_exact"] - so["h_iters_exact"])
        worst_p, worst_h = max(worst_p, dp), max(worst_h, dh)
        assert dp <= 1 and dh <= 1, (step, st, so)
        if step % 100 == 99:
            print(f"substep {step + 1}: pressure its gpu {st['p_iters_exact']} oracle {so['p_iters_exact']}, "
                  f"{time.time() - t0:.0f} s", flush=True)
    return sim, orc, worst_p, worst_h


def test_plume_500_substeps_track_the_oracle(gpu):
    kw, dt, _ = plume_scene()
    f64 = O.OracleSim(**kw, precision="f64")
    sim, orc, wp, wh = _lockstep(kw, dt, K, f64)
    print("worst CG count difference: pressure", wp, "heat", wh, flush=True)
    for f in ("T", "U", "V", "W", "P"):
        g, o, t = sim._scene.get_field(f), orc.field(f), f64.field(f)
        err = rel_l2(g, o)
        print(f"{f}: gpu vs fp32 oracle {err:.2e}; vs fp64 oracle: gpu {rel_l2(g, t):.2e}, "
              f"fp32 oracle {rel_l2(o, t):.2e}", flush=True)
        if f != "P":
            assert err <= 1e-3, f
    sim.close()


def test_coupled_n64_500_substeps_track_the_oracle(gpu):
    kw, dt, _ = coupled_scene(64)
    sim, orc, wp, wh = _lockstep(kw, dt, K)
    print("particles", sim.n, "worst CG count difference: pressure", wp, "heat", wh, flush=True)
    for f in ("T", "U", "V", "W"):
        err = rel_l2(sim._scene.get_field(f), orc.field(f))
        print(f, "rel_l2 after 500 substeps", err, flush=True)
        assert err <= 1e-3, f
    err = rel_l2(sim.temperatures(), orc.temperatures())
    print("particle temperatures rel_l2", err, flush=True)
    assert err <= 1e-3
    sim.close()
