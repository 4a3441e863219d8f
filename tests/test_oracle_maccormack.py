"""Limited MacCormack advection in the oracle (HFM_FLAG_MACCORMACK), pinned by what the method
guarantees rather than by a copy of its formula:
* the limiter: no value leaves the range of the start field (a maximum principle);
* second order: a smooth bump keeps more of its peak than under the semi-Lagrangian step, and
  moves by the same distance;
* exactness: an affine field under a uniform velocity is reproduced exactly away from the walls,
  as by the semi-Lagrangian step (trilinear interpolation is exact for affine fields, so the
  correction is zero), and a uniform face velocity stays uniform away from the inflow wall."""
import numpy as np
import pytest

import oracle as O
from heatflow_mpm import GridConfig, Heater, _abi

NONE = np.zeros((0, 3), np.float32)
N = 24
H = 0.01
U0 = 0.6            # m/s along x: 0.12 cells per 2 ms step
DT = 2e-3


def _sim(flags, precision="f64"):
    return O.OracleSim(NONE, heater=Heater(on=False), grid=GridConfig(L=N * H, h=H), precision=precision,
                       flags=flags)


def _uniform_u(sim):
    u = np.zeros((N + 1, N, N))
    u[1:N] = U0
    sim.set_field("U", u)
    sim.set_field("V", np.zeros((N, N + 1, N)))
    sim.set_field("W", np.zeros((N, N, N + 1)))


def _advect_T(sim, T0, K):
    T = T0.copy()
    for _ in range(K):
        sim.set_field("TW", T)
        sim.stage("ADVECT_T", DT)
        T = sim.field("T2")
    return T


def _bump():
    c = (np.arange(N) + 0.5) * H
    x, y, z = np.meshgrid(c, c, c, indexing="ij")
    return 20.0 + 30.0 * np.exp(-((x - 0.08) ** 2 + (y - 0.12) ** 2 + (z - 0.12) ** 2) / (2 * 0.02 ** 2))


@pytest.fixture(scope="module")
def bumps():
    T0 = _bump()
    out = {}
    for name, flags in (("sl", 0), ("mc", _abi.FLAG_MACCORMACK)):
        sim = _sim(flags)
        _uniform_u(sim)
        out[name] = _advect_T(sim, T0, 60)
    return T0, out


def test_limiter_keeps_the_range(bumps):
    T0, out = bumps
    for T in out.values():
        assert T.max() <= T0.max() + 1e-12 and T.min() >= T0.min() - 1e-12


def test_maccormack_diffuses_less(bumps):
    T0, out = bumps
    peak0 = T0.max() - 20.0
    sl, mc = out["sl"].max() - 20.0, out["mc"].max() - 20.0
    print(f"peak after 60 steps: semi-Lagrangian {sl / peak0:.3f}, MacCormack {mc / peak0:.3f} of the start")
    assert mc > sl + 0.05 * peak0
    # both move the bump by U dt K (0.072 m): the centroid of the excess along x
    c = (np.arange(N) + 0.5) * H
    for T in out.values():
        ex = T - 20.0
        cx = (ex.sum(axis=(1, 2)) * c).sum() / ex.sum()
        assert cx == pytest.approx(0.08 + U0 * DT * 60, abs=0.004)


def test_affine_field_is_exact():
    c = (np.arange(N) + 0.5) * H
    x, y, z = np.meshgrid(c, c, c, indexing="ij")
    T0 = 10.0 + 200.0 * x - 50.0 * y + 80.0 * z
    exact = T0 - 200.0 * U0 * DT
    inner = (slice(3, N - 3),) * 3
    for flags in (0, _abi.FLAG_MACCORMACK):
        sim = _sim(flags)
        _uniform_u(sim)
        T = _advect_T(sim, T0, 1)
        assert np.max(np.abs(T[inner] - exact[inner])) < 1e-10


def test_uniform_face_velocity_stays_uniform():
    for flags in (0, _abi.FLAG_MACCORMACK):
        sim = _sim(flags)
        _uniform_u(sim)
        sim.stage("ADVECT_FACES", DT)
        u = sim.field("U")
        # the first interior face traces back toward the wall face (u = 0); from the second on the
        # departure points and the forward traces fall between faces that hold U0
        assert np.allclose(u[2:N], U0, rtol=0, atol=1e-12) and np.all(u[0] == 0) and np.all(u[N] == 0)
        assert np.all(sim.field("V") == 0) and np.all(sim.field("W") == 0)


def test_face_bump_keeps_its_range_and_peak():
    c = (np.arange(N) + 0.5) * H
    x, y, z = np.meshgrid(c, (np.arange(N + 1)) * H, c, indexing="ij")
    v0 = 0.2 * np.exp(-((x - 0.08) ** 2 + (y - 0.12) ** 2 + (z - 0.12) ** 2) / (2 * 0.02 ** 2))
    v0[:, 0] = v0[:, N] = 0.0
    peaks = {}
    for name, flags in (("sl", 0), ("mc", _abi.FLAG_MACCORMACK)):
        sim = _sim(flags)
        v = v0.copy()
        for _ in range(40):
            _uniform_u(sim)
            sim.set_field("V", v)
            sim.stage("ADVECT_FACES", DT)
            v = sim.field("V")
        assert v.max() <= v0.max() + 1e-12 and v.min() >= -1e-12
        peaks[name] = v.max()
    assert peaks["mc"] > peaks["sl"]
