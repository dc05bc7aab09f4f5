"""Full-size checks at the netflix-shaped workload (480,189 x 17,770,
99,072,112 train ratings, f = 100): the shape the benchmark runs.

The FP64 oracle cannot redo whole half-sweeps here in test time, so parity is
checked on sampled rows from the GPU's own state (uniform users, and the
heaviest items, which the schedule splits across waves), plus the test RMSE
recomputed on the host from the downloaded factors.
Tolerances (relative to max(1, |x|)): mixed path (FP16 Gram, FP32 Cholesky,
one FP64-residual refinement step, classic or warm-started) 5e-5 against the
FP64 solve of the FP16-rounded system: rows whose residual is formed from the
restaged ratings land within ~3e-8, rows refined against the FP32-accumulated
tiles (long and split rows, every warm-started row, and the light rows solved
through their dual system) keep that accumulation's error, ~1e-5 on the
heaviest items.
FP32 path 1e-4 against the exact FP64 solve.
"""
import numpy as np
import pytest

import cuda_als_amd as ca
from cuda_als_amd import engine as E
from oracle import als_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def netflix():
    return ca.netflix_shape().generate()


@pytest.mark.parametrize("prec,refine,warm,oprec,tol", [(E.Prec.FP16, 1, False, O.PREC_FP16, 5e-5),
                                                        (E.Prec.FP16, 1, True, O.PREC_FP16, 5e-5),
                                                        (E.Prec.FP32, 0, False, O.PREC_EXACT, 1e-4)])
def test_sampled_rows_match_oracle_at_full_size(netflix, prec, refine, warm, oprec, tol):
    d, f = netflix, 100
    rng = np.random.default_rng(11)
    cnt_u = np.diff(d.csr_ptr)
    # uniform users plus users light enough for the dual system
    users = np.unique(np.concatenate([rng.choice(d.m, 300, replace=False),
                                      rng.choice(np.where(cnt_u < 48)[0], 100, replace=False)]))
    cnt = np.diff(d.csc_ptr)
    items = np.sort(np.concatenate([np.argsort(cnt)[-40:], rng.choice(d.n, 60, replace=False)]))
    with E.Engine(E.make_config(f, 0.05, 1, E.SolveType.CHOL, prec, refine=refine, refine_warm=warm), d) as e:
        e.init_factors()
        e.sweep()
        e.sweep()
        V = e.get_factors(ca.ALS_SIDE_ITEMS).astype(np.float64)
        e.update(ca.ALS_SIDE_USERS)
        U1 = e.get_factors(ca.ALS_SIDE_USERS)
        e.update(ca.ALS_SIDE_ITEMS)
        V1 = e.get_factors(ca.ALS_SIDE_ITEMS)
        info = e.info()
        rmse = e.rmse()
    assert info.failures == 0
    assert info.nsplit_rows[1] > 0   # heavy items really were split
    cfg = O.config(f, 0.05, O.SOLVE_CHOL, gram_prec=oprec)
    Uo = O.update_rows(cfg, users, d.csr_ptr, d.csr_idx, d.csr_val, V)
    eu = np.abs(U1[users] - Uo).max(axis=1) / np.maximum(1.0, np.abs(Uo).max(axis=1))
    Vo = O.update_rows(cfg, items, d.csc_ptr, d.csc_idx, d.csc_val, U1.astype(np.float64))
    ei = np.abs(V1[items] - Vo).max(axis=1) / np.maximum(1.0, np.abs(Vo).max(axis=1))
    print("users max rel err", eu.max(), "items max rel err", ei.max())
    assert eu.max() <= tol and ei.max() <= tol
    # test RMSE: FP64 recomputation from the factors
    pred = np.einsum("ij,ij->i", U1[d.t_row].astype(np.float64), V1[d.t_col].astype(np.float64))
    ref = np.sqrt(np.mean((d.t_val - pred) ** 2))
    assert abs(rmse - ref) <= 1e-6 * ref
