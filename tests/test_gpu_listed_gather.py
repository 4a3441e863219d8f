"""The cell lattice's particle sum of a body that stays put comes from a contribution list built
once (the gather's terms in the gather's order); it must give the same bits as the general gather,
which a KinematicBody with zero velocity runs every substep (its particles are uploaded and
re-binned each time)."""
import numpy as np
import pytest

from heatflow_mpm import HeatflowBatch, HeatflowMPM, KinematicBody, coupled_scene

pytestmark = pytest.mark.gpu
FIELDS = ("T", "U", "V", "W", "P", "TSUM", "ALPHA_S")


def still(pos, t):
    return np.zeros_like(pos)


@pytest.mark.parametrize("n", [40, 64])
def test_listed_gather_equals_the_general_gather(gpu, n):
    kw, dt, _ = coupled_scene(n)
    listed = HeatflowMPM(**kw)
    general = HeatflowMPM(**kw, body=KinematicBody(kw["particles"], still))
    for step in range(6):
        listed.substep(dt)
        general.substep(dt)
        for f in FIELDS:
            assert np.array_equal(listed._scene.get_field(f), general._scene.get_field(f)), (step, f)
        assert np.array_equal(listed.temperatures(), general.temperatures()), step
    # new particle positions invalidate the list: the next substeps re-bin, count and rebuild
    moved = kw["particles"] + np.float32(0.004)
    listed.body.pos[:] = moved
    listed._scene.update_particles(listed.body.pos, listed.body.vel)
    general.body.pos[:] = moved
    for step in range(4):
        listed.substep(dt)
        general.substep(dt)
        for f in FIELDS:
            assert np.array_equal(listed._scene.get_field(f), general._scene.get_field(f)), ("moved", step, f)


def test_listed_gather_in_a_batch(gpu):
    """Scenes at different points of the list's life in one batch: the batch keeps each scene's bits."""
    kws = [coupled_scene(m)[0] for m in (32, 40, 48)]
    dt = coupled_scene(32)[1]
    alone = []
    for kw in kws:
        sim = HeatflowMPM(**kw)
        for _ in range(5):
            sim.substep(dt)
        alone.append(sim._scene.get_field("T"))
        sim.close()
    sims = [HeatflowMPM(**kw) for kw in kws]
    sims[1].substep(dt)            # one scene is a substep ahead when the batch forms
    sims[1].substep(dt)
    with HeatflowBatch(sims) as b:
        b.run(dt, 3)
    sims[0].substep(dt)
    sims[0].substep(dt)
    sims[2].substep(dt)
    sims[2].substep(dt)
    for sim, ref in zip(sims, alone):
        assert np.array_equal(sim._scene.get_field("T"), ref)
