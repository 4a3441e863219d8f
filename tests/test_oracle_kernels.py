"""Pins the CPU oracle's kernels against independent constructions (CPU only).

Each check is chosen so that a plausible slip in the restatement (a dropped term, a wrong sign
or index, a transposed operand) fails: numpy float32 reproductions of the fp32 op sequences,
scipy's linear interpolation and sparse matrices built from the operators' definitions, closed
forms and conservation identities.
"""
import numpy as np
import pytest
import scipy.ndimage as ndi
import scipy.sparse as sp

import oracle as O

OFFS = {"cell": (0.5, 0.5, 0.5), "x": (0.0, 0.5, 0.5), "y": (0.5, 0.0, 0.5), "z": (0.5, 0.5, 0.0)}


# ---- fields.py:5-48 ------------------------------------------------------------------------

def test_base_index_is_the_fp32_op_sequence():
    """fields.py:5-10: g = pos/h - off in fp32, i = floor(g), f = g - i — bit for bit."""
    rng = np.random.default_rng(3)
    h = np.float32(0.005)
    pos = rng.uniform(-0.01, 0.49, (2000, 3)).astype(np.float32)
    # include exact lattice points and half points, where rounding decides the bin
    pos[:200] = (rng.integers(0, 96, (200, 3)) * 0.005).astype(np.float32)
    pos[200:400] = ((rng.integers(0, 96, (200, 3)) + 0.5) * 0.005).astype(np.float32)
    for off in OFFS.values():
        o = np.asarray(off, np.float32)
        g = pos / h - o
        i_ref = np.floor(g).astype(np.int32)
        f_ref = (g - i_ref.astype(np.float32)).astype(np.float32)
        for t in range(0, len(pos), 7):
            bi, f = O.base_index(pos[t].astype(np.float64), np.asarray(off), float(h))
            assert np.array_equal(bi, i_ref[t]), (t, off)
            assert np.array_equal(np.float32(f), f_ref[t]), (t, off)


@pytest.mark.parametrize("lattice", ["cell", "x", "z"])
def test_sample_equals_clamped_linear_interpolation(lattice):
    """fields.py:25-36 = trilinear interpolation with coordinates clamped to the lattice
    (scipy map_coordinates order 1, mode 'nearest'), including points outside the box."""
    rng = np.random.default_rng(5)
    n, h = 5, 0.1
    dims = [n + (lattice == "x"), n, n + (lattice == "z")]
    a = rng.standard_normal(dims)
    off = np.asarray(OFFS[lattice])
    for _ in range(60):
        pos = rng.uniform(-0.15, n * h + 0.15, 3)
        coord = np.clip(pos / h - off, 0, np.asarray(dims) - 1)
        ref = ndi.map_coordinates(a, coord.reshape(3, 1), order=1, mode="nearest")[0]
        got = O.sample(a, pos, off, h, precision="f64")
        assert abs(got - ref) < 1e-12


def test_sample_reproduces_nodes_and_affine_fields():
    n, h = 6, 0.05
    i, j, k = np.meshgrid(*[np.arange(n)] * 3, indexing="ij")
    a = 1.5 * i - 2.0 * j + 0.25 * k + 3.0          # affine in lattice index
    off = np.asarray(OFFS["cell"])
    rng = np.random.default_rng(1)
    for _ in range(40):
        c = rng.uniform(0.0, n - 1, 3)               # inside the node hull
        pos = (c + off) * h
        assert abs(O.sample(a, pos, off, h, precision="f64") - (1.5 * c[0] - 2 * c[1] + 0.25 * c[2] + 3)) < 1e-11
    node = np.array([2, 3, 4])
    assert O.sample(a, (node + off) * h, off, h, precision="f64") == pytest.approx(a[2, 3, 4], abs=1e-12)


def test_scatter_conserves_mass_and_is_the_adjoint_of_sample():
    """fields.py:39-48: sum(scatter(v)) = v even when clamped; <scatter(v), a> = v * sample(a)."""
    rng = np.random.default_rng(2)
    n, h = 5, 0.1
    a = rng.standard_normal((n, n + 1, n))
    off = np.asarray(OFFS["y"])
    for _ in range(30):
        pos = rng.uniform(-0.1, n * h + 0.1, 3)
        v = rng.uniform(-2, 2)
        s = O.scatter(np.zeros_like(a), pos, off, h, v, precision="f64")
        assert s.sum() == pytest.approx(v, abs=1e-12)
        assert (s * a).sum() == pytest.approx(v * O.sample(a, pos, off, h, precision="f64"), abs=1e-12)


def test_backtrace_of_uniform_flow_is_exact():
    """fields.py:111-116: with uniform velocity U the departure point is x - dt U."""
    n, h, dt = 6, 0.04, 0.01
    U = np.array([0.7, -0.4, 1.1])
    u = np.full((n + 1, n, n), U[0])
    v = np.full((n, n + 1, n), U[1])
    w = np.full((n, n, n + 1), U[2])
    x = np.array([0.11, 0.13, 0.09])
    assert np.allclose(O.backtrace(u, v, w, n, h, dt, x, precision="f64"), x - dt * U, atol=1e-15)


# ---- pressure operator, grid.py:94-137 --------------------------------------------------------

def _face_fractions(rng, n):
    ax = rng.uniform(0.05, 1.0, (n + 1, n, n))
    ay = rng.uniform(0.05, 1.0, (n, n + 1, n))
    az = rng.uniform(0.05, 1.0, (n, n, n + 1))
    ax[[0, -1]] = 0
    ay[:, [0, -1]] = 0
    az[:, :, [0, -1]] = 0
    return ax, ay, az


def _pressure_matrix_from_definition(n, h, ax, ay, az):
    """sum over interior faces f = (c, nb) of a_f (e_c - e_nb)(e_c - e_nb)^T / h^2"""
    idx = np.arange(n ** 3).reshape(n, n, n)
    rows, cols, vals = [], [], []

    def add(c, d, a):
        for (p, q, s) in ((c, c, a), (d, d, a), (c, d, -a), (d, c, -a)):
            rows.append(p), cols.append(q), vals.append(s / h ** 2)

    for i in range(1, n):
        for j in range(n):
            for k in range(n):
                add(idx[i - 1, j, k], idx[i, j, k], ax[i, j, k])
                add(idx[j, i - 1, k], idx[j, i, k], ay[j, i, k])
                add(idx[j, k, i - 1], idx[j, k, i], az[j, k, i])
    return sp.csr_matrix((vals, (rows, cols)), shape=(n ** 3, n ** 3)).toarray()


def test_laplacian_matches_its_definition():
    rng = np.random.default_rng(11)
    n, h = 4, 0.01
    ax, ay, az = _face_fractions(rng, n)
    N = n ** 3
    A = np.column_stack([O.laplacian_apply(n, h, ax, ay, az, e, precision="f64") for e in np.eye(N)])
    ref = _pressure_matrix_from_definition(n, h, ax, ay, az)
    assert np.allclose(A, ref, rtol=1e-12, atol=1e-9)
    assert np.allclose(A, A.T)
    assert np.abs(A @ np.ones(N)).max() < 1e-8                  # constants in the null space
    assert np.linalg.eigvalsh(A).min() > -1e-6                 # PSD
    d = O.jacobi_diag(n, h, ax, ay, az, precision="f64")
    assert np.allclose(1.0 / d, np.diag(A), rtol=1e-12)        # Jacobi = inverse diagonal
    # alpha / beta form: out = alpha A x + beta y
    x, y = rng.standard_normal(N), rng.standard_normal(N)
    out = O.laplacian_apply(n, h, ax, ay, az, x, y, alpha=-1.0, beta=1.0, precision="f64")
    assert np.allclose(out, y - A @ x, rtol=1e-12, atol=1e-8)


# ---- heat operator, heat.py:50-144 -------------------------------------------------------------

def _heat_matrix_from_definition(n, h, dt, kap, fixed):
    N = n ** 3
    A = np.eye(N)
    idx = np.arange(N).reshape(n, n, n)
    c = dt / h ** 2
    for i in range(n):
        for j in range(n):
            for k in range(n):
                t = idx[i, j, k]
                if fixed[i, j, k]:
                    continue
                for d in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
                    ii, jj, kk = i + d[0], j + d[1], k + d[2]
                    if not (0 <= ii < n and 0 <= jj < n and 0 <= kk < n):
                        continue
                    kf = 2 * kap[i, j, k] * kap[ii, jj, kk] / (kap[i, j, k] + kap[ii, jj, kk] + 1e-12)
                    A[t, t] += c * kf
                    if not fixed[ii, jj, kk]:
                        A[t, idx[ii, jj, kk]] -= c * kf
    return A


def test_heat_operator_rhs_and_diagonal_match_their_definition():
    rng = np.random.default_rng(4)
    n, h, dt = 4, 0.01, 0.5
    kap = rng.uniform(1e-5, 5e-5, (n, n, n))
    fixed = (rng.uniform(size=(n, n, n)) < 0.25).astype(np.int32)
    tfix = rng.uniform(20, 80, (n, n, n))
    N = n ** 3
    A = np.column_stack([O.heat_apply(n, h, dt, kap, fixed, tfix, e, precision="f64") for e in np.eye(N)])
    ref = _heat_matrix_from_definition(n, h, dt, kap, fixed)
    assert np.allclose(A, ref, rtol=1e-12, atol=1e-12)
    free = fixed.ravel() == 0
    Aff = A[np.ix_(free, free)]
    assert np.allclose(Aff, Aff.T) and np.linalg.eigvalsh(Aff).min() > 0   # SPD on the free cells
    T = rng.uniform(20, 30, (n, n, n))
    rhs, diag = O.heat_rhs_diag(n, h, dt, T, kap, fixed, tfix, precision="f64")
    # rhs: fixed cells carry T_fix; free cells T + dt/h^2 sum over fixed neighbours kappa_f T_fix,
    # i.e. moving the fixed-neighbour columns of the full backward-Euler operator to the right side
    full = _heat_matrix_from_definition(n, h, dt, kap, np.zeros_like(fixed))
    Tf = np.where(fixed.ravel() == 1, tfix.ravel(), 0.0)
    offdiag = full - np.diag(np.diag(full))          # -dt/h^2 kappa_f between neighbours
    exp = np.where(fixed.ravel() == 1, tfix.ravel(), T.ravel() - offdiag @ Tf)
    assert np.allclose(rhs, exp, rtol=1e-12)
    # diagonal preconditioner: 1 / (1 + dt/h^2 sum_nb kappa_f) including fixed neighbours
    assert np.allclose(1.0 / diag, np.where(free, np.diag(full), 1.0), rtol=1e-12)


def test_heat_operator_uniform_kappa_is_the_textbook_stencil():
    n, h, dt, kv = 5, 0.02, 0.3, 2e-4
    kap = np.full((n, n, n), kv)
    fixed = np.zeros((n, n, n), np.int32)
    x = np.random.default_rng(0).standard_normal((n, n, n))
    out = O.heat_apply(n, h, dt, kap, fixed, np.zeros_like(kap), x, precision="f64").reshape(n, n, n)
    kf = 2 * kv * kv / (2 * kv + 1e-12)
    lap = np.zeros_like(x)
    for ax in range(3):
        for s in (1, -1):
            sh = np.roll(x, s, axis=ax)
            edge = [slice(None)] * 3
            edge[ax] = 0 if s == 1 else -1
            diff = x - sh
            diff[tuple(edge)] = 0.0                  # no neighbour across the box boundary
            lap += diff
    assert np.allclose(out, x + dt / h ** 2 * kf * lap, rtol=1e-12, atol=1e-12)


# ---- PCG driver (warp.optim.linear.cg semantics) ------------------------------------------------

def test_pcg_solves_the_heat_system():
    rng = np.random.default_rng(8)
    n, h, dt = 4, 0.01, 2.0
    kap = rng.uniform(1e-5, 5e-5, (n, n, n))
    fixed = (rng.uniform(size=(n, n, n)) < 0.2).astype(np.int32)
    tfix = rng.uniform(20, 80, (n, n, n))
    T = rng.uniform(20, 30, (n, n, n))
    rhs, diag = O.heat_rhs_diag(n, h, dt, T, kap, fixed, tfix, precision="f64")
    A = _heat_matrix_from_definition(n, h, dt, kap, fixed)
    x, exact, reported, conv, rr, atol = O.pcg("heat", n, h, diag, rhs, T, 1e-10, 200, 10, dt=dt, kap=kap,
                                               fixed=fixed, tfix=tfix, precision="f64")
    assert conv and rr <= atol ** 2
    assert np.allclose(x, np.linalg.solve(A, rhs), rtol=1e-8)
    assert reported == 10 * -(-exact // 10)


def test_pcg_solves_the_pressure_system_and_matches_scipy_iterations():
    import scipy.sparse.linalg as spla
    rng = np.random.default_rng(9)
    n, h = 5, 0.01
    ax, ay, az = _face_fractions(rng, n)
    A = _pressure_matrix_from_definition(n, h, ax, ay, az)
    b = rng.standard_normal(n ** 3)
    b -= b.mean()                                   # consistent right-hand side (grid.py:238-241)
    d = O.jacobi_diag(n, h, ax, ay, az, precision="f64")
    tol = 1e-8
    x, exact, reported, conv, rr, atol = O.pcg("pressure", n, h, d, b, np.zeros(n ** 3), tol, 400, 20,
                                               af=(ax, ay, az), precision="f64")
    assert conv
    assert np.linalg.norm(A @ x - b) <= 1.0001 * atol
    assert atol == pytest.approx(max(tol * np.linalg.norm(b), tol))
    its = []
    spla.cg(A, b, x0=np.zeros_like(b), rtol=0.0, atol=atol, maxiter=400, M=sp.diags(d),
            callback=lambda xk: its.append(1))
    assert abs(len(its) - exact) <= 1


@pytest.mark.parametrize("exact,converged,maxiter,ce,expected", [
    (1, True, 200, 10, 10), (10, True, 200, 10, 10), (11, True, 200, 10, 20), (23, True, 400, 20, 40),
    (20, True, 400, 20, 20), (395, True, 400, 20, 400), (400, False, 400, 20, 400), (7, False, 7, 20, 8),
    (199, True, 200, 10, 200)])
def test_warp_iteration_count_quantisation(exact, converged, maxiter, ce, expected):
    """warp's _run_solver_loop: cycles of 2, a host check when count % check_every < 2, and a stop
    once the count reaches maxiter."""
    assert O.warp_reported_count(exact, converged, maxiter, ce) == expected


# ---- numpy float32 mean (grid.py:239) ------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 7, 8, 127, 128, 129, 8191, 8192, 8193, 110592, 216000, 884736])
def test_rhs_mean_is_numpys_float32_mean_bit_for_bit(n):
    rng = np.random.default_rng(n)
    a = (rng.standard_normal(n) * rng.uniform(0.01, 1e3)).astype(np.float32)
    assert O.numpy_mean_f32(a).tobytes() == a.mean().tobytes()
    assert O.numpy_sum_f32(a).tobytes() == np.add.reduce(a).tobytes()
