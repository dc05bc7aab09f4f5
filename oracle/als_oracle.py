"""ctypes binding of the FP64 oracle (oracle/als_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, always as the checker or the baseline, never as
part of the product path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libals_oracle.so"
SOLVE_LU, SOLVE_CG_FP16A, SOLVE_CG, SOLVE_CHOL, SOLVE_CG32 = 0, 1, 2, 3, 4
PREC_EXACT, PREC_BF16, PREC_FP16, PREC_FP32 = 0, 1, 2, 3


class OrcConfig(C.Structure):
    _fields_ = [("f", C.c_int), ("lambda_", C.c_double), ("solver", C.c_int), ("cg_iters", C.c_int),
                ("cg_tol", C.c_double), ("gram_prec", C.c_int), ("store_fp32", C.c_int)]


P = C.c_void_p
_SIG = {
    "orc_reference_init": (None, [C.c_int64, C.c_int64, C.c_int, P, P]),
    "orc_round": (C.c_double, [C.c_double, C.c_int]),
    "orc_gram_rhs": (None, [C.c_int, C.c_int64, P, P, P, C.c_double, C.c_int, P, P]),
    "orc_chol_solve": (C.c_int, [C.c_int, P, P]),
    "orc_lu_nopiv_solve": (C.c_int, [C.c_int, P, P]),
    "orc_cg_solve": (C.c_int, [C.c_int, P, P, P, C.c_int, C.c_double]),
    "orc_cg_solve_f32": (C.c_int, [C.c_int, P, P, P, C.c_int, C.c_double]),
    "orc_update": (C.c_int64, [C.POINTER(OrcConfig), C.c_int64, P, P, P, P, P, P]),
    "orc_update_rows": (C.c_int64, [C.POINTER(OrcConfig), C.c_int64, P, P, P, P, P, P, P, P]),
    "orc_rmse": (C.c_double, [C.c_int64, P, P, P, P, P, C.c_int]),
    "orc_objective": (C.c_double, [C.c_int, C.c_double, C.c_int64, P, P, P, C.c_int64, P, P, P]),
    "orc_train": (C.c_int64, [C.POINTER(OrcConfig), C.c_int64, C.c_int64, P, P, P, P, P, P, C.c_int64, P, P, P,
                              C.c_int, P, P, P]),
    "orc_adjoint": (None, [C.c_int, C.c_int64, C.c_int64, P, P, P, P, P, P, P, P, P, P]),
    "orc_num_threads": (C.c_int, []),
    "orc_set_num_threads": (None, [C.c_int]),
}


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} missing: python cuda_als_amd/build.py --oracle")
    lib = C.CDLL(str(LIB_PATH))
    for k, (r, a) in _SIG.items():
        fn = getattr(lib, k)
        fn.restype, fn.argtypes = r, a
    return lib


lib = _load()


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def config(f, lambda_=0.05, solver=SOLVE_CHOL, cg_iters=6, cg_tol=1e-4, gram_prec=PREC_EXACT, store_fp32=False):
    return OrcConfig(f, lambda_, solver, cg_iters, cg_tol, gram_prec, int(store_fp32))


def reference_init(m, n, f):
    U = np.zeros((m, f)); V = np.zeros((n, f))
    lib.orc_reference_init(m, n, f, _p(U), _p(V))
    return U, V


def round_prec(x: float, prec: int) -> float:
    return lib.orc_round(float(x), prec)


def gram_rhs(f, idx, val, Y, lambda_, prec=PREC_EXACT):
    idx, val, Y = _c(idx, np.int32), _c(val, np.float32), _c(Y, np.float64)
    A = np.zeros((f, f)); b = np.zeros(f)
    lib.orc_gram_rhs(f, idx.size, _p(idx), _p(val), _p(Y), lambda_, prec, _p(A), _p(b))
    return A, b


def chol_solve(A, b):
    A = _c(A, np.float64).copy(); b = _c(b, np.float64).copy()
    rc = lib.orc_chol_solve(A.shape[0], _p(A), _p(b))
    return rc, b


def lu_solve(A, b):
    A = _c(A, np.float64).copy(); b = _c(b, np.float64).copy()
    rc = lib.orc_lu_nopiv_solve(A.shape[0], _p(A), _p(b))
    return rc, b


def cg_solve(A, b, x0, iters, tol, fp32=False):
    A = _c(A, np.float64); b = _c(b, np.float64); x = _c(x0, np.float64).copy()
    (lib.orc_cg_solve_f32 if fp32 else lib.orc_cg_solve)(A.shape[0], _p(A), _p(b), _p(x), iters, tol)
    return x


def update(cfg, ptr, idx, val, Y, X, lambda_rows=None):
    """One half-sweep; returns (X_new, failures)."""
    ptr, idx, val = _c(ptr, np.int32), _c(idx, np.int32), _c(val, np.float32)
    Y = _c(Y, np.float64); X = _c(X, np.float64).copy()
    lr = None if lambda_rows is None else _c(lambda_rows, np.float64)
    bad = lib.orc_update(C.byref(cfg), ptr.size - 1, _p(ptr), _p(idx), _p(val), _p(Y), _p(X), _p(lr))
    return X, bad


def update_rows(cfg, sel, ptr, idx, val, Y, X_in=None, lambda_rows=None):
    sel = _c(sel, np.int64); ptr, idx, val = _c(ptr, np.int32), _c(idx, np.int32), _c(val, np.float32)
    Y = _c(Y, np.float64)
    out = np.zeros((sel.size, cfg.f))
    xin = None if X_in is None else _c(X_in, np.float64)
    lr = None if lambda_rows is None else _c(lambda_rows, np.float64)
    lib.orc_update_rows(C.byref(cfg), sel.size, _p(sel), _p(ptr), _p(idx), _p(val), _p(Y), _p(out), _p(xin), _p(lr))
    return out


def rmse(row, col, val, U, V):
    row, col, val = _c(row, np.int32), _c(col, np.int32), _c(val, np.float32)
    U, V = _c(U, np.float64), _c(V, np.float64)
    return lib.orc_rmse(row.size, _p(row), _p(col), _p(val), _p(U), _p(V), U.shape[1])


def objective(lambda_, d, U, V):
    U, V = _c(U, np.float64), _c(V, np.float64)
    return lib.orc_objective(U.shape[1], lambda_, d.m, _p(d.csr_ptr), _p(d.csr_idx), _p(d.csr_val), d.n,
                             _p(d.csc_ptr), _p(U), _p(V))


def train(cfg, d, sweeps, U0, V0):
    """Returns (U, V, rmse_per_sweep, failures)."""
    U = _c(U0, np.float64).copy(); V = _c(V0, np.float64).copy()
    r = np.zeros(sweeps)
    bad = lib.orc_train(C.byref(cfg), d.m, d.n, _p(d.csr_ptr), _p(d.csr_idx), _p(d.csr_val), _p(d.csc_ptr),
                        _p(d.csc_idx), _p(d.csc_val), d.n_test, _p(d.t_row), _p(d.t_col), _p(d.t_val), sweeps,
                        _p(U), _p(V), _p(r))
    return U, V, r, bad


def adjoint(f, ptr, idx, val, Y, lambda_rows, X, Gbar):
    """Returns (Xbar, lambda_bar, Ybar) of the half-sweep X = solve(A(lambda, Y), b(Y))."""
    ptr, idx, val = _c(ptr, np.int32), _c(idx, np.int32), _c(val, np.float32)
    Y, X, G = _c(Y, np.float64), _c(X, np.float64), _c(Gbar, np.float64)
    lr = _c(lambda_rows, np.float64)
    rows, ncols = ptr.size - 1, Y.shape[0]
    Xb = np.zeros((rows, f)); lb = np.zeros(rows); Yb = np.zeros((ncols, f))
    lib.orc_adjoint(f, rows, ncols, _p(ptr), _p(idx), _p(val), _p(Y), _p(lr), _p(X), _p(G), _p(Xb), _p(lb), _p(Yb))
    return Xb, lb, Yb
