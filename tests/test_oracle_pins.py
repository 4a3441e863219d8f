"""Closed-form pins of the oracle functions that the other CPU tests leave open (CPU only).

Face fraction (grid.py:46-59), the conductivity ramp (heat.py:44-47), the diagnostics
cell_velocity / divergence_max / speed_max (grid.py:156-162, 253-263), the PCG's breakdown stop and
the heater energy balance (heater.py:52, heat.py:154-158).  The expected values are worked out by
hand from the definitions (not by re-running the formula), and each case is chosen so that a
plausible slip fails: a swapped neighbour, a missing 0.5, a clamp edge taken on the wrong side, a
wall face not zeroed, a transposed axis, an af index one off, a sign of the flux term.
"""
import numpy as np
import pytest

import oracle as O
from heatflow_mpm import GridConfig, Heater, Material
from heatflow_mpm.solver import KAPPA_AIR

NONE = np.zeros((0, 3), np.float32)
N8 = 8


def _box(precision="f32", mat=Material(), cg_tol=1e-6):
    """An 8^3 box of 1 cm cells; the heater's default position lies outside it, so no cell is solid,
    fixed walls aside (the alpha and face fields below are set directly)."""
    return O.OracleSim(NONE, mat=mat, heater=Heater(power=0.0), grid=GridConfig(L=0.08, h=0.01, cg_tol=cg_tol),
                       precision=precision)


# (alpha of the lower cell, alpha of the upper cell) -> air fraction of the face between them, by hand:
# clamp(1 - (a + b) / 2, 0.05, 1)
FACE_CASES = [((0.0, 0.0), 1.0), ((1.0, 1.0), 0.05), ((0.3, 0.9), 0.4), ((1.2, 0.9), 0.05),
              ((0.2, 0.2), 0.8), ((0.02, 0.0), 0.99), ((0.0, 1.0), 0.5), ((1.9, 0.0), 0.05),
              ((0.9, 0.9), 0.1), ((0.95, 0.95), 0.05), ((0.0, -0.2), 1.0)]


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_face_fraction_hand_values_and_wall_faces(axis):
    sim = _box()
    n = N8
    alpha = np.zeros((n, n, n))
    # case c sits on line c of the two transverse axes, between cells 3 and 4 of `axis`
    for c, ((a, b), _) in enumerate(FACE_CASES):
        p, q = c // n, c % n
        lo = [p, q]
        lo.insert(axis, 3)
        hi = [p, q]
        hi.insert(axis, 4)
        alpha[tuple(lo)] = a
        alpha[tuple(hi)] = b
    sim.set_field("ALPHA_S", alpha)
    sim.stage("FACE_FRACTION")
    af = sim.field(["AX", "AY", "AZ"][axis])
    for c, ((a, b), want) in enumerate(FACE_CASES):
        p, q = c // n, c % n
        f = [p, q]
        f.insert(axis, 4)          # face 4 lies between cells 3 and 4
        assert af[tuple(f)] == pytest.approx(want, abs=1e-6), (a, b)
    wall0 = [slice(None)] * 3
    wall0[axis] = 0
    walln = [slice(None)] * 3
    walln[axis] = n
    assert np.all(af[tuple(wall0)] == 0.0) and np.all(af[tuple(walln)] == 0.0)
    # faces away from the planted pairs see alpha = 0 on both sides
    far = [7, 7]
    far.insert(axis, 2)
    assert af[tuple(far)] == 1.0


def test_conductivity_ramp_hand_values():
    """kappa = kappa_air + (kappa_mat - kappa_air) clamp(alpha / 0.6, 0, 1), with kappa_air 2.2e-5,
    kappa_mat 1e-6: values worked out by hand at the ramp's ends, middle and outside it."""
    sim = _box(mat=Material(kappa=1.0e-6))
    assert KAPPA_AIR == 2.2e-5
    cases = {0.0: 2.2e-5, 0.15: 1.675e-5, 0.3: 1.15e-5, 0.45: 6.25e-6, 0.6: 1.0e-6, 1.2: 1.0e-6, -0.3: 2.2e-5}
    alpha = np.zeros((N8, N8, N8))
    flat = alpha.reshape(-1)
    for c, a in enumerate(cases):
        flat[c] = a
    sim.set_field("ALPHA_S", alpha)
    sim.stage("CONDUCTIVITY")
    kap = sim.field("KAPPA").reshape(-1)
    for c, want in enumerate(cases.values()):
        assert kap[c] == pytest.approx(want, rel=2e-6), list(cases)[c]


def _set_faces(sim, u=None, v=None, w=None, alpha=None):
    n = N8
    sim.set_field("ALPHA_S", np.zeros((n, n, n)) if alpha is None else alpha)
    sim.stage("FACE_FRACTION")
    for name, val, shape in (("U", u, (n + 1, n, n)), ("V", v, (n, n + 1, n)), ("W", w, (n, n, n + 1))):
        sim.set_field(name, np.zeros(shape) if val is None else val)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_divergence_and_speed_max_of_a_linear_field(axis):
    """Face velocity c * (face index) along one axis, solid-free box: interior cells have div = c / h,
    the last cell meets the wall face (air fraction 0), so div = -c (n - 1) / h there; hence
    divergence_max = c (n - 1) and speed_max = c n (the wall face's own value)."""
    c = 0.25
    n = N8
    sim = _box(precision="f64")
    idx = np.arange(n + 1, dtype=float)
    shape = [1, 1, 1]
    shape[axis] = n + 1
    full = [n, n, n]
    full[axis] = n + 1
    field = np.broadcast_to((c * idx).reshape(shape), full).copy()
    kw = {["u", "v", "w"][axis]: field}
    _set_faces(sim, **kw)
    dm, sm = sim.diagnostics()
    assert dm == pytest.approx(c * (n - 1), rel=1e-12)
    assert sm == pytest.approx(c * n, rel=1e-12)


def test_divergence_includes_the_particle_flux_with_its_sign():
    """q enters as div(a_f u + q): u = 0 and q = 0.5 on the x faces 1..n-1 (zero on the walls) gives
    div = +0.5 / h in cell 0 and -0.5 / h in cell n - 1, 0 between: divergence_max = 0.5."""
    n = N8
    sim = _box(precision="f64")
    _set_faces(sim)
    q = np.zeros((n + 1, n, n))
    q[1:n] = 0.5
    sim.set_field("QX", q)
    dm, sm = sim.diagnostics()
    assert dm == pytest.approx(0.5, rel=1e-12)
    assert sm == 0.0
    q[1:n] = 0.5 + 0.1 * np.arange(1, n).reshape(-1, 1, 1)      # interior cells now see +0.1 / h
    sim.set_field("QX", q)
    dm, _ = sim.diagnostics()
    assert dm == pytest.approx(0.5 + 0.1 * (n - 1), rel=1e-12)  # the last cell: -(q[n-1] - 0) / h


def test_cell_velocity_of_an_affine_field_is_the_field_at_the_cell_centres():
    """Faces carry u = 1 + 2x - y + 0.5z (and v, w other affine fields) at their own centres; the face
    average of grid.py:156-162 is then the field at the cell centre (exact for affine fields)."""
    n, h = N8, 0.01
    sim = _box(precision="f64")

    def at(offs, dims, coef):
        i, j, k = np.meshgrid(*[np.arange(d, dtype=float) for d in dims], indexing="ij")
        x, y, z = ((i + offs[0]) * h, (j + offs[1]) * h, (k + offs[2]) * h)
        return coef[0] + coef[1] * x + coef[2] * y + coef[3] * z

    cu, cv, cw = (1.0, 2.0, -1.0, 0.5), (-0.5, 0.0, 3.0, 1.0), (0.25, -2.0, 0.0, 4.0)
    _set_faces(sim, u=at((0, 0.5, 0.5), (n + 1, n, n), cu), v=at((0.5, 0, 0.5), (n, n + 1, n), cv),
               w=at((0.5, 0.5, 0), (n, n, n + 1), cw))
    cell = sim.air_velocity().reshape(n, n, n, 3)
    for d, cf in enumerate((cu, cv, cw)):
        want = at((0.5, 0.5, 0.5), (n, n, n), cf)
        assert np.allclose(cell[..., d], want, rtol=0, atol=1e-12), d


def test_pcg_breakdown_stops_and_keeps_the_last_iterate():
    """An indefinite system (negative face coefficients on half the faces) makes p.Ap non-positive
    within a few iterations: the solve stops there with converged = -1 and a finite x, where warp's
    loop would run on with a negative or infinite step.  A definite system of the same size
    converges with converged = 1 (the stop changes nothing there)."""
    n, h = 6, 0.1
    rng = np.random.default_rng(0)
    ax, ay, az = (np.ones((n + 1, n, n)), np.ones((n, n + 1, n)), np.ones((n, n, n + 1)))
    b = rng.standard_normal(n ** 3)
    b -= b.mean()
    diag = O.jacobi_diag(n, h, ax, ay, az, precision="f64")
    x, it, _, conv, rr, atol = O.pcg("pressure", n, h, diag, b, np.zeros(n ** 3), 1e-8, 400, 20, af=(ax, ay, az),
                                     precision="f64")
    assert conv and np.isfinite(x).all()
    bad = [a.copy() for a in (ax, ay, az)]
    for a in bad:
        a[..., ::2] *= -1.0
    x, it, _, conv, rr, atol = O.pcg("pressure", n, h, np.abs(diag), b, np.zeros(n ** 3), 1e-8, 400, 20,
                                     af=tuple(bad), precision="f64")
    assert it < 400 and np.isfinite(x).all() and np.isfinite(rr)
    assert O.lib("f64").orc_real_bytes() == 8
    # the stop is reported as a breakdown, not as convergence or the cap
    st = _breakdown_flag(n, h, bad, b, diag)
    assert st == -1


def _breakdown_flag(n, h, af, b, diag):
    import ctypes as C
    x = np.zeros(n ** 3)
    iters = np.zeros(3, np.int32)
    rr = np.zeros(2)
    keep = [np.ascontiguousarray(a, np.float64).ravel() for a in af]
    O.lib("f64").orc_pcg(0, n, h, 0.0, *[k.ctypes.data_as(C.c_void_p) for k in keep], None, None, None,
                         np.ascontiguousarray(np.abs(diag)), np.ascontiguousarray(b), x, 1e-8, 400, 20, iters, rr)
    return int(iters[2])


def test_source_energy_balance_in_the_stats():
    """heater.py:52 sets rate = P / (rho c_p N_src h^3), so the fp64 sum of the source step's rises
    times rho c_p h^3 is P dt while no source cell reaches its cap (heat.py:154-158)."""
    dt = 2e-3
    sim = O.OracleSim(NONE, heater=Heater(power=400.0, t_cap=120.0), grid=GridConfig(L=0.48, h=0.01), precision="f32")
    sim.substep(dt)
    rise = sim.stats()["source_dT_sum"]
    assert rise * 1.2 * 1005.0 * 0.01 ** 3 == pytest.approx(400.0 * dt, rel=1e-5)
    off = O.OracleSim(NONE, heater=Heater(power=0.0), grid=GridConfig(L=0.48, h=0.01), precision="f32")
    off.substep(dt)
    assert off.stats()["source_dT_sum"] == 0.0
