"""Moving material (solver.py:59,83-88 with a body that moves): particles with nonzero velocities
feed the flux q and relaxation target m of the air (solver.py:19-38), are re-binned after every
move, and receive temperatures, melt fractions, yield stresses and viscosities back.  A
KinematicBody (prescribed swirl and rise) drives the GPU scene and, with an identical copy, the
oracle; the two must agree like the static scenes do."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2
from heatflow_mpm import HeatflowMPM, KinematicBody, coupled_scene

pytestmark = pytest.mark.gpu
FIELDS = ["T", "U", "V", "W", "P", "ALPHA_S", "ALPHA_P", "TSUM", "QX", "QY", "QZ", "KX", "KY", "KZ", "MX", "MY", "MZ"]


def swirl(pos, t):
    """Rigid rotation at 2 rad/s about the vertical axis through the box centre, rising at 2 cm/s."""
    r = pos[:, :2] - np.float32(0.24)
    v = np.empty_like(pos)
    v[:, 0] = -2.0 * r[:, 1]
    v[:, 1] = 2.0 * r[:, 0]
    v[:, 2] = 0.02
    return v.astype(np.float32)


def test_moving_body_matches_the_oracle(gpu):
    kw, dt, _ = coupled_scene(40)
    P = kw["particles"]
    sim = HeatflowMPM(**kw, body=KinematicBody(P, swirl))
    body_o = KinematicBody(P, swirl)
    orc = O.OracleSim(**kw, velocities=body_o.vel)
    for step in range(4):
        sim.substep(dt)
        orc.update_particles(body_o.pos, body_o.vel)
        orc.substep(dt)
        body_o.step(dt)
        st, so = sim.stats(), orc.stats()
        print("substep", step, "pressure its", st["p_iters_exact"], so["p_iters_exact"], flush=True)
        assert abs(st["p_iters_exact"] - so["p_iters_exact"]) <= 1
    assert np.array_equal(sim.positions(), body_o.pos)
    for f in FIELDS:
        err = rel_l2(sim._scene.get_field(f), orc.field(f))
        print(f, "rel_l2", err, flush=True)
        assert err <= 1e-5, f
    got = [sim.temperatures(), sim.melt_fraction(), sim.yield_stress(), sim.viscosity()]
    for name, g, o in zip(("temp", "melt", "yield", "visc"), got, orc.particles()):
        err = rel_l2(g, o)
        print(name, "rel_l2", err, flush=True)
        assert err <= 1e-5, name
    # the material moved: its flux and relaxation targets are not zero
    assert np.abs(sim._scene.get_field("QX")).max() > 0 and np.abs(sim._scene.get_field("MZ")).max() > 0
