"""Limited MacCormack advection on the GPU (HFM_FLAG_MACCORMACK) against the oracle with the same
flag: the plain parity bars (1e-5 after one substep, CG counts within 1, 1e-4 after 20
substeps); a MacCormack scene gives the same bits alone and in a batch with a semi-Lagrangian
one."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2
from heatflow_mpm import HeatflowBatch, HeatflowMPM, _abi, coupled_scene, plume_scene

pytestmark = pytest.mark.gpu
FIELDS = ["T", "U", "V", "W", "P", "KAPPA"]
MC = _abi.FLAG_MACCORMACK


@pytest.mark.parametrize("scene", ["plume48", "coupled40"])
def test_maccormack_matches_the_oracle(gpu, scene):
    kw, dt, _ = plume_scene() if scene == "plume48" else coupled_scene(40)
    sim = HeatflowMPM(**kw, advection="maccormack")
    orc = O.OracleSim(**kw, flags=MC)
    sim.substep(dt)
    orc.substep(dt)
    st, so = sim.stats(), orc.stats()
    assert abs(st["p_iters_exact"] - so["p_iters_exact"]) <= 1 and abs(st["h_iters_exact"] - so["h_iters_exact"]) <= 1
    for f in FIELDS:
        err = rel_l2(sim._scene.get_field(f), orc.field(f))
        print(scene, f, "rel_l2", err, flush=True)
        assert err <= 1e-5, f
    if scene == "plume48":
        for _ in range(19):
            sim.substep(dt)
            orc.substep(dt)
        for f in ("T", "U", "V", "W"):
            err = rel_l2(sim._scene.get_field(f), orc.field(f))
            print("20 substeps", f, "rel_l2", err, flush=True)
            assert err <= 1e-4, f
    else:
        assert rel_l2(sim.temperatures(), orc.temperatures()) <= 1e-5


def test_maccormack_differs_from_semi_lagrangian(gpu):
    kw, dt, _ = plume_scene()
    a = HeatflowMPM(**kw, advection="maccormack")
    b = HeatflowMPM(**kw)
    for _ in range(20):
        a.substep(dt)
        b.substep(dt)
    d = rel_l2(a._scene.get_field("T"), b._scene.get_field("T"))
    assert 1e-6 < d < 1e-1


def test_maccormack_alone_equals_in_a_batch(gpu):
    kws = [plume_scene()[0], coupled_scene(40)[0]]
    dt = plume_scene()[1]
    modes = ["maccormack", "semi-lagrangian"]
    alone = []
    for kw, m in zip(kws, modes):
        sim = HeatflowMPM(**kw, advection=m)
        for _ in range(3):
            sim.substep(dt)
        alone.append({f: sim._scene.get_field(f) for f in FIELDS})
        sim.close()
    sims = [HeatflowMPM(**kw, advection=m) for kw, m in zip(kws, modes)]
    with HeatflowBatch(sims) as b:
        b.run(dt, 3)
    for sim, ref in zip(sims, alone):
        for f, v in ref.items():
            assert np.array_equal(sim._scene.get_field(f), v), f
