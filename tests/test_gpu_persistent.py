"""The persistent pressure PCG (HFM_FLAG_PERSISTENT_CG): every iteration of the one-launch form in one
cooperative launch per solve, each workgroup keeping its tile's vectors on chip.  It restates the
one-launch form on the same tiles with the same partial sums and reduction order, so it is held to
equality with it: the same bits in every field and the same iteration counts, convergence flags and
residuals, substep after substep.  The one-launch form is held to the fp32 oracle elsewhere; one
substep of the persistent form is checked against the oracle here as well.  Every case asserts that the
persistent kernel really ran (its kernel-timing class counts one launch per pressure solve), or, where
the scene cannot run it, that it did not."""
import threading

import numpy as np
import pytest

import oracle as O
from conftest import rel_l2
from heatflow_mpm import HeatflowBatch, HeatflowMPM, coupled_scene, plume_scene
from heatflow_mpm import _abi

pytestmark = pytest.mark.gpu
FIELDS = ["T", "U", "V", "W", "P", "RHS"]
STATS = ["p_iters", "p_iters_exact", "p_converged", "p_rr", "h_iters_exact", "source_dT_sum"]


def _sim(kw, cg, planes):
    sim = HeatflowMPM(**kw, cg=cg, layout=dict(dir_planes=planes))
    sim._scene.set_kernel_timing(True)
    return sim


def _persist_launches(sim):
    return sim._scene.kernel_time(_abi.KT_PCG_PERSIST)[1]


def _assert_same(a, b, where):
    sa, sb = a.stats(), b.stats()
    for k in STATS:
        assert sa[k] == sb[k], (where, k, sa[k], sb[k])
    for f in FIELDS:
        x, y = a._scene.get_field(f), b._scene.get_field(f)
        assert np.array_equal(x, y), (where, f, float(np.abs(x - y).max()))


def _lockstep(kw, dt, substeps, planes=16, expect_persistent=True):
    ref, per = _sim(kw, "fused", planes), _sim(kw, "persistent", planes)
    try:
        for k in range(substeps):
            ref.substep(dt)
            per.substep(dt)
            _assert_same(ref, per, f"substep {k + 1}")
        st = per.stats()
        print("n", kw["grid"].n, "planes", planes, "p_iters_exact", st["p_iters_exact"], "persistent launches",
              _persist_launches(per), flush=True)
        assert _persist_launches(per) == (substeps if expect_persistent else 0)
        assert _persist_launches(ref) == 0
    finally:
        ref.close()
        per.close()


def test_persistent_coupled_96_equals_one_launch_form(gpu):
    """216 tiles of 16^3, one per CU; the first substep from rest runs to the 400-iteration cap."""
    kw, dt, _ = coupled_scene(96)
    _lockstep(kw, dt, 4)


def test_persistent_plume_many_substeps_equal_one_launch_form(gpu):
    kw, dt, _ = plume_scene()
    _lockstep(kw, dt, 40)


def test_persistent_ragged_tiles_and_shallow_blocks(gpu):
    """n = 40: tiles cut by the box on every axis, with 16-, 8- and 4-plane blocks (27, 45, 90 tiles)."""
    kw, dt, _ = coupled_scene(40)
    for planes in (16, 8, 4):
        _lockstep(kw, dt, 6, planes=planes)


def test_persistent_one_substep_matches_oracle(gpu):
    kw, dt, _ = plume_scene()
    sim, orc = _sim(kw, "persistent", 16), O.OracleSim(**kw)
    sim.substep(dt)
    orc.substep(dt)
    st, so = sim.stats(), orc.stats()
    print("persistent", st["p_iters_exact"], "oracle", so["p_iters_exact"], flush=True)
    assert _persist_launches(sim) == 1
    assert abs(st["p_iters_exact"] - so["p_iters_exact"]) <= 1
    for f in FIELDS:
        err = rel_l2(sim._scene.get_field(f), orc.field(f))
        print(f, "rel_l2", err, flush=True)
        assert err <= 1e-5, f
    sim.close()


def test_persistent_falls_back_where_it_cannot_run(gpu):
    """A batch of two runs the persistent form in one cooperative launch (same bits as each scene stepped
    alone); at n = 128 the 512 tiles do not fit one per CU and the scene runs the one-launch form."""
    kw, dt, _ = coupled_scene(48)
    alone = [_sim(kw, "persistent", 16) for _ in range(2)]
    pair = [_sim(kw, "persistent", 16) for _ in range(2)]
    try:
        with HeatflowBatch(pair) as batch:
            for k in range(5):
                batch.substep(dt)
                for s in alone:
                    s.substep(dt)
                for a, b in zip(alone, pair):
                    _assert_same(a, b, f"batch substep {k + 1}")
        assert all(_persist_launches(s) == 5 for s in alone)
    finally:
        for s in alone + pair:
            s.close()
    kw, dt, _ = coupled_scene(128)
    _lockstep(kw, dt, 2, expect_persistent=False)


def _solo_fused_run(kw, dt, substeps, planes=16):
    sim = _sim(kw, "fused", planes)
    for _ in range(substeps):
        sim.substep(dt)
    return sim


def test_persistent_batch_equals_scenes_alone(gpu):
    """A batch of scenes with the persistent form: those whose tiles fit the device (n = 32, 40 with ragged
    tiles, 48, 64) solve in one cooperative launch, each scene's workgroups waiting only on each other; the
    n = 128 scene (512 tiles) runs the one-launch form in the same batch.  Every scene ends with the bits of
    the one-launch form stepped alone."""
    from heatflow_mpm.scenes import random_scene
    rng = np.random.default_rng(7)
    kws = [random_scene(rng, n, name=f"b{i}")[1] for i, n in enumerate((32, 40, 48, 64, 48, 128))]
    dt, K = 2e-3, 4
    sims = [_sim(kw, "persistent", 16) for kw in kws]
    try:
        with HeatflowBatch(sims) as batch:
            batch.set_kernel_timing(True)
            for _ in range(K):
                batch.substep(dt)
            groups = batch.kernel_time(_abi.KT_PCG_PERSIST)[1]
        print("persistent group launches", groups, flush=True)
        assert groups == K          # the five fitting scenes share one launch per substep
        for i, (kw, s) in enumerate(zip(kws, sims)):
            ref = _solo_fused_run(kw, dt, K)
            _assert_same(ref, s, f"scene {i} (n = {kw['grid'].n})")
            ref.close()
    finally:
        for s in sims:
            s.close()


def test_persistent_batch_of_large_scenes_takes_several_launches(gpu):
    """Three 96^3 scenes (216 workgroups each) do not fit one launch: one group each, in turn."""
    from heatflow_mpm.scenes import random_scene
    rng = np.random.default_rng(11)
    kws = [random_scene(rng, 96, name=f"l{i}")[1] for i in range(3)]
    dt, K = 2e-3, 2
    sims = [_sim(kw, "persistent", 16) for kw in kws]
    try:
        with HeatflowBatch(sims) as batch:
            batch.set_kernel_timing(True)
            for _ in range(K):
                batch.substep(dt)
            assert batch.kernel_time(_abi.KT_PCG_PERSIST)[1] == 3 * K
        for i, (kw, s) in enumerate(zip(kws, sims)):
            ref = _solo_fused_run(kw, dt, K)
            _assert_same(ref, s, f"scene {i}")
            ref.close()
    finally:
        for s in sims:
            s.close()


def test_persistent_wait_timeout_is_reported(gpu, monkeypatch):
    """The bounded wait's failure path, forced from the host: with HFM_PCG_PERSIST_SPINS=-1 every workgroup
    gives up at its first wait for the others, reports the solve as stopped and ends; the substep returns
    the error.  Afterwards, with the limit back, scenes solve as before (same bits as the one-launch form)."""
    kw, dt, _ = coupled_scene(48)   # from rest the fan drives a nonzero divergence: the solve iterates
    sim = _sim(kw, "persistent", 16)
    monkeypatch.setenv("HFM_PCG_PERSIST_SPINS", "-1")
    with pytest.raises(RuntimeError, match="ran out of polls"):
        sim.substep(dt)
    assert _persist_launches(sim) == 1
    sim.close()
    monkeypatch.delenv("HFM_PCG_PERSIST_SPINS")
    _lockstep(kw, dt, 2)


def test_persistent_solves_from_two_threads(gpu):
    """Two scenes stepped from two host threads at once (the library runs their persistent solves one
    at a time on the device): each ends with the bits of the same scene stepped alone."""
    kw, dt, _ = plume_scene()
    ref = _sim(kw, "persistent", 16)
    for _ in range(12):
        ref.substep(dt)
    sims = [_sim(kw, "persistent", 16) for _ in range(2)]
    errors = []

    def run(s):
        try:
            for _ in range(12):
                s.substep(dt)
            s.synchronize()
        except Exception as e:   # reported below, in the test's thread
            errors.append(e)

    threads = [threading.Thread(target=run, args=(s,)) for s in sims]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for s in sims:
        _assert_same(ref, s, "threaded")
        assert _persist_launches(s) == 12
    for s in sims + [ref]:
        s.close()
