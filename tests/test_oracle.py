"""Pins the FP64 oracle (CPU only).

Each function of oracle/als_oracle.c is checked against something that does
not restate its own formula: glibc's documented rand() sequence, numpy/scipy
routines, brute-force dense computation on tiny inputs, closed forms,
first-order optimality of the half-step, monotone descent of the ALS
objective, and central finite differences for the adjoint.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from oracle import als_oracle as O


# --------------------------------------------------------------- init rule
def test_reference_init_uses_glibc_rand_sequence():
    """als_model.cu:1772-1777: srand(0) (== seed 1 in glibc) then
    VT[k] = 0.2 * ((float)rand() / (float)RAND_MAX). glibc's first outputs for
    seed 1 are 1804289383, 846930886, 1681692777, 1714636915, 1957747793."""
    U, V = O.reference_init(3, 2, 3)
    first = [1804289383, 846930886, 1681692777, 1714636915, 1957747793]
    rand_max = np.float32(2147483647)          # (float)RAND_MAX rounds to 2^31
    want = [np.float32(0.2 * np.float64(np.float32(np.float32(r) / rand_max))) for r in first]
    assert np.array_equal(V.ravel()[:5].astype(np.float32), np.array(want, np.float32))
    assert not U.any()


# --------------------------------------------------------------- rounding
def test_round_fp16_matches_numpy():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(2000) * s for s in (1e-6, 1e-3, 1.0, 30.0)])
    x = np.concatenate([x, [6.1e-5, 5.96e-8, 3e-8, 65504.0, -0.0]])
    got = np.array([O.round_prec(v, O.PREC_FP16) for v in x])
    want = x.astype(np.float32).astype(np.float16).astype(np.float64)
    assert np.array_equal(got, want)


def test_round_bf16_matches_bit_rounding():
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(4000) * rng.choice([1e-3, 1.0, 1e3], 4000)).astype(np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16        # round to nearest even on the top 16 bits
    want = r.astype(np.uint32).view(np.float32).astype(np.float64)
    got = np.array([O.round_prec(float(v), O.PREC_BF16) for v in x])
    assert np.array_equal(got, want)


# ------------------------------------------------------------ Gram + RHS
def test_gram_rhs_brute_force_dense():
    """A_u = Y^T diag(mask_u) Y + lambda n_u I and b_u = Y^T r_u computed from a
    dense masked rating row (a different route than the per-rating sum)."""
    rng = np.random.default_rng(2)
    n_items, f, lam = 9, 5, 0.07
    Y = rng.standard_normal((n_items, f))
    idx = np.array([0, 3, 4, 8], np.int32)
    val = np.array([1, 5, 3, 2], np.float32)
    A, b = O.gram_rhs(f, idx, val, Y, lam)
    mask = np.zeros(n_items); mask[idx] = 1
    r = np.zeros(n_items); r[idx] = val
    assert np.allclose(A, Y.T @ np.diag(mask) @ Y + lam * idx.size * np.eye(f), atol=1e-12)
    assert np.allclose(b, Y.T @ r, atol=1e-12)


def test_gram_invariants():
    rng = np.random.default_rng(3)
    Y = rng.standard_normal((40, 12))
    idx = rng.choice(40, 6, replace=False).astype(np.int32)
    lam = 0.05
    A, _ = O.gram_rhs(12, idx, np.ones(6, np.float32), Y, lam)
    assert np.allclose(A, A.T)
    ev = np.linalg.eigvalsh(A - lam * 6 * np.eye(12))
    assert np.sum(ev > 1e-9) == 6                      # rank <= n_u
    assert np.all(np.linalg.eigvalsh(A) > 0)           # SPD
    assert np.isclose(np.trace(A), np.sum(Y[idx] ** 2) + 12 * lam * 6)


# ---------------------------------------------------------------- solvers
def test_cholesky_and_lu_match_library():
    rng = np.random.default_rng(4)
    M = rng.standard_normal((30, 30))
    A = M @ M.T + 30 * np.eye(30)
    b = rng.standard_normal(30)
    want = sl.cho_solve(sl.cho_factor(A), b)
    rc1, x1 = O.chol_solve(A, b)
    rc2, x2 = O.lu_solve(A, b)
    assert rc1 == 0 and rc2 == 0
    assert np.allclose(x1, want, rtol=1e-12, atol=1e-12) and np.allclose(x2, want, rtol=1e-11, atol=1e-11)
    assert np.allclose(x2, np.linalg.solve(A, b))


def test_cholesky_reports_non_spd():
    rc, _ = O.chol_solve(np.array([[1.0, 2.0], [2.0, 1.0]]), np.ones(2))
    assert rc != 0


def test_cg_with_enough_iterations_is_exact():
    rng = np.random.default_rng(5)
    M = rng.standard_normal((12, 12))
    A = M @ M.T + 5 * np.eye(12)
    b = rng.standard_normal(12)
    x = O.cg_solve(A, b, np.zeros(12), 40, 0.0)
    assert np.allclose(x, np.linalg.solve(A, b), atol=1e-10)


def test_cg_one_by_one_converges_in_one_step():
    x = O.cg_solve(np.array([[4.0]]), np.array([2.0]), np.array([7.0]), 1, 0.0)
    assert np.isclose(x[0], 0.5)


def test_cg_early_exit_after_first_update():
    """cumf_als_cg/cg.cu:196: the r.r < tol test comes after x is updated, so
    with a large tol exactly one step is taken: x1 = x0 + a r0, a = r0.r0 / r0.A r0."""
    rng = np.random.default_rng(6)
    M = rng.standard_normal((6, 6))
    A = M @ M.T + np.eye(6)
    b, x0 = rng.standard_normal(6), rng.standard_normal(6)
    r0 = b - A @ x0
    want = x0 + (r0 @ r0) / (r0 @ A @ r0) * r0
    assert np.allclose(O.cg_solve(A, b, x0, 6, 1e30), want, atol=1e-12)


def test_cg_warm_start_at_solution_stays_finite():
    """p.Ap == 0 (r == 0) stops the recurrence instead of dividing 0/0."""
    A = np.diag([2.0, 3.0])
    x = np.array([0.5, 1.0])
    out = O.cg_solve(A, A @ x, x, 6, 0.0)
    assert np.all(np.isfinite(out)) and np.allclose(out, x)


def test_cg32_is_the_fp32_execution_of_the_recurrence():
    """ORC_SOLVE_CG32: every iterate is float32; with enough iterations it
    solves to FP32 accuracy (numpy's solve); the first step equals the
    closed form x0 + (r.r / r.Ar) r to FP32 rounding; a 1 x 1 system
    converges in one step; on a well-conditioned system it stays within a
    few float32 ulps of the FP64 recurrence after each of 1..4 iterations."""
    rng = np.random.default_rng(8)
    M = rng.standard_normal((12, 12))
    A = M @ M.T + 12 * np.eye(12)
    b = rng.standard_normal(12)
    x = O.cg_solve(A, b, np.zeros(12), 40, 0.0, fp32=True)
    assert np.array_equal(x, x.astype(np.float32).astype(np.float64))
    want = np.linalg.solve(A, b)
    assert 0 < np.abs(x - want).max() <= 2e-5 * np.abs(want).max()
    x0 = rng.standard_normal(12)
    r0 = b - A @ x0
    one = x0 + (r0 @ r0) / (r0 @ A @ r0) * r0
    assert np.abs(O.cg_solve(A, b, x0, 6, 1e30, fp32=True) - one).max() <= 1e-5 * np.abs(one).max()
    assert abs(O.cg_solve(np.array([[4.0]]), np.array([2.0]), np.array([7.0]), 1, 0.0, fp32=True)[0] - 0.5) < 1e-7
    for it in range(1, 5):
        x64 = O.cg_solve(A, b, x0, it, 0.0)
        x32 = O.cg_solve(A, b, x0, it, 0.0, fp32=True)
        assert np.abs(x32 - x64).max() <= 1e-5 * np.abs(x64).max(), it


# ------------------------------------------------------------- half-sweep
def _tiny_data(seed=7, m=30, n=20, density=0.3, empty_row=None):
    import cuda_als_amd as ca
    rng = np.random.default_rng(seed)
    dense = (rng.random((m, n)) < density) * rng.integers(1, 6, (m, n))
    if empty_row is not None:
        dense[empty_row] = 0
    rows, cols = np.nonzero(dense)
    ptr = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=m))]).astype(np.int32)
    return ca.from_arrays(m, n, ptr, cols, dense[rows, cols], rows[:10], cols[:10], dense[rows[:10], cols[:10]])


def test_half_step_zeroes_the_gradient():
    """After the exact update, d/dx_u [sum (r - x_u.y_i)^2 + lambda n_u |x_u|^2] = 0,
    evaluated from the raw ratings."""
    d = _tiny_data()
    f, lam = 4, 0.1
    rng = np.random.default_rng(8)
    V = rng.standard_normal((d.n, f))
    U, bad = O.update(O.config(f, lam, O.SOLVE_CHOL), d.csr_ptr, d.csr_idx, d.csr_val, V, np.zeros((d.m, f)))
    assert bad == 0
    for u in range(d.m):
        s, e = d.csr_ptr[u], d.csr_ptr[u + 1]
        y = V[d.csr_idx[s:e]]
        g = -2 * y.T @ (d.csr_val[s:e] - y @ U[u]) + 2 * lam * (e - s) * U[u]
        assert np.abs(g).max() < 1e-10


def test_empty_row_gets_zero():
    d = _tiny_data(empty_row=3)
    U, _ = O.update(O.config(4, 0.1, O.SOLVE_CHOL), d.csr_ptr, d.csr_idx, d.csr_val,
                    np.ones((d.n, 4)), np.full((d.m, 4), 9.0))
    assert not U[3].any()


def test_objective_is_monotone_over_half_steps():
    d = _tiny_data(seed=9, m=40, n=25)
    f, lam = 5, 0.05
    cfg = O.config(f, lam, O.SOLVE_CHOL)
    U, V = O.reference_init(d.m, d.n, f)
    last = np.inf
    for _ in range(6):
        U, _ = O.update(cfg, d.csr_ptr, d.csr_idx, d.csr_val, V, U)
        o1 = O.objective(lam, d, U, V)
        V, _ = O.update(cfg, d.csc_ptr, d.csc_idx, d.csc_val, U, V)
        o2 = O.objective(lam, d, U, V)
        assert o1 <= last + 1e-9 and o2 <= o1 + 1e-9
        last = o2


def test_f1_reduces_to_scalar_ridge():
    """f = 1: x_u = sum r y / (sum y^2 + lambda n_u)."""
    d = _tiny_data(seed=10)
    lam = 0.3
    V = np.random.default_rng(11).random((d.n, 1)) + 0.5
    U, _ = O.update(O.config(1, lam, O.SOLVE_LU), d.csr_ptr, d.csr_idx, d.csr_val, V, np.zeros((d.m, 1)))
    for u in range(d.m):
        s, e = d.csr_ptr[u], d.csr_ptr[u + 1]
        y = V[d.csr_idx[s:e], 0]
        want = (d.csr_val[s:e] @ y) / (y @ y + lam * (e - s)) if e > s else 0.0
        assert np.isclose(U[u, 0], want, rtol=1e-12)


def test_rmse_closed_form():
    rng = np.random.default_rng(12)
    U, V = rng.standard_normal((5, 3)), rng.standard_normal((4, 3))
    row = np.array([0, 1, 4, 2], np.int32); col = np.array([3, 0, 1, 1], np.int32)
    val = np.array([1.0, 2.0, 3.5, 4.0], np.float32)
    pred = np.einsum("ij,ij->i", U[row], V[col])
    assert np.isclose(O.rmse(row, col, val, U, V), np.sqrt(np.mean((val - pred) ** 2)), rtol=1e-14)


def test_train_matches_manual_half_steps():
    d = _tiny_data(seed=13)
    f, lam = 3, 0.05
    cfg = O.config(f, lam, O.SOLVE_CHOL)
    U0, V0 = O.reference_init(d.m, d.n, f)
    U, V, r, bad = O.train(cfg, d, 2, U0, V0)
    Um, Vm = U0, V0
    for _ in range(2):
        Um, _ = O.update(cfg, d.csr_ptr, d.csr_idx, d.csr_val, Vm, Um)
        Vm, _ = O.update(cfg, d.csc_ptr, d.csc_idx, d.csc_val, Um, Vm)
    assert np.allclose(U, Um) and np.allclose(V, Vm)
    assert np.isclose(r[-1], O.rmse(d.t_row, d.t_col, d.t_val, U, V))


# ---------------------------------------------------------------- adjoint
def test_adjoint_matches_central_differences():
    """L = sum_u g_u . x_u(lambda, Y) for a fixed g: dL/dlambda_u and dL/dY from
    the oracle adjoint against FP64 central differences (h = 1e-6)."""
    d = _tiny_data(seed=14, m=12, n=10, density=0.5)
    f = 3
    rng = np.random.default_rng(15)
    Y = rng.standard_normal((d.n, f))
    lam = 0.05 + rng.random(d.m) * 0.1
    G = rng.standard_normal((d.m, f))
    cfg = O.config(f, 0.0, O.SOLVE_CHOL)

    def loss(lam_, Y_):
        X, _ = O.update(cfg, d.csr_ptr, d.csr_idx, d.csr_val, Y_, np.zeros((d.m, f)), lam_)
        return float(np.sum(G * X))

    X, _ = O.update(cfg, d.csr_ptr, d.csr_idx, d.csr_val, Y, np.zeros((d.m, f)), lam)
    Xb, lb, Yb = O.adjoint(f, d.csr_ptr, d.csr_idx, d.csr_val, Y, lam, X, G)
    h = 1e-6
    for u in range(d.m):
        e = np.zeros(d.m); e[u] = h
        fd = (loss(lam + e, Y) - loss(lam - e, Y)) / (2 * h)
        assert abs(fd - lb[u]) <= 1e-6 * max(1.0, abs(fd))
    for i in range(d.n):
        for j in range(f):
            E = np.zeros_like(Y); E[i, j] = h
            fd = (loss(lam, Y + E) - loss(lam, Y - E)) / (2 * h)
            assert abs(fd - Yb[i, j]) <= 1e-6 * max(1.0, abs(fd))
