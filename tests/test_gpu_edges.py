"""Grid sizes at the edge of what the library accepts, on the GPU against the fp32 oracle.

The smallest grids and odd sizes (n = 2, 3, 5, 7, 13): every PCG tile, workgroup and lattice row is cut
by the box, a 2-cell row is the shortest one a trilinear sample reads as one pair, and the pressure system
has a handful of unknowns.  Particles hotter than the air sit inside the box (their heat drives a buoyant
flow), the heater is off and the walls are fixed.  Each pressure-PCG form is run for three substeps and
held to the oracle's bars: iteration counts within one, T, u, v, w and p within relative L2 1e-5, with an
absolute floor of 1e-9 for a field that stays at rounding level.  n = 1 is rejected with an error."""
import numpy as np
import pytest

import oracle as O
from heatflow_mpm import GridConfig, Heater, HeatflowMPM, Material

pytestmark = pytest.mark.gpu


def _scene(n):
    L = 0.01 * n
    rng = np.random.default_rng(n)
    P = rng.uniform(0.25 * L, 0.75 * L, size=(4 * n * n, 3)).astype(np.float32)
    return dict(particles=P, mat=Material(t_sol=28.0, t_liq=34.0), heater=Heater(on=False),
                grid=GridConfig(L=L, h=0.01, cg_tol=1e-6), spacing=2.5e-3, t_init=60.0)


def _close(g, o):
    g = np.asarray(g, np.float64).ravel()
    o = np.asarray(o, np.float64).ravel()
    d = float(np.linalg.norm(g - o))
    den = float(np.linalg.norm(o))
    return d <= 1e-5 * den or d <= 1e-9, (d / den if den > 0 else d)


@pytest.mark.parametrize("cg", ["warp", "fused", "persistent"])
@pytest.mark.parametrize("n", [2, 3, 5, 7, 13])
def test_small_and_odd_grids_match_the_oracle(gpu, n, cg):
    kw = _scene(n)
    dt = 2e-3
    sim = HeatflowMPM(**kw, cg=cg)
    orc = O.OracleSim(**kw)
    try:
        for step in range(3):
            sim.substep(dt)
            orc.substep(dt)
            st, so = sim.stats(), orc.stats()
            assert abs(st["p_iters_exact"] - so["p_iters_exact"]) <= 1, (step, st["p_iters_exact"], so["p_iters_exact"])
            assert abs(st["h_iters_exact"] - so["h_iters_exact"]) <= 1, (step, st["h_iters_exact"], so["h_iters_exact"])
        worst = 0.0
        for f in ("T", "U", "V", "W", "P"):
            ok, err = _close(sim._scene.get_field(f), orc.field(f))
            worst = max(worst, err)
            assert ok, (f, err)
        ok, err = _close(sim.temperatures(), orc.temperatures())
        assert ok, ("particle temperatures", err)
        print(f"n={n} cg={cg} layout={sim.layout}: pressure its {st['p_iters_exact']} (oracle {so['p_iters_exact']}), "
              f"heat its {st['h_iters_exact']}, worst field rel L2 {worst:.2e}, speed_max {sim.grid.speed_max():.3e}",
              flush=True)
        assert np.isfinite(sim.grid_T()).all()
    finally:
        sim.close()


def test_one_cell_grid_is_rejected(gpu):
    with pytest.raises(RuntimeError):
        HeatflowMPM(np.zeros((0, 3), np.float32), heater=Heater(on=False), grid=GridConfig(L=0.01, h=0.01))
