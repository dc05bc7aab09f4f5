"""GPU parity: the HIP engine (through its C ABI) against the FP64 oracle on
identical seeded inputs.

Tolerances (absolute):
  FP32 path: per-sweep test RMSE and factors within 1e-4 of the oracle.
  Mixed precision (FP16 or BF16 Gram, FP32 Cholesky + FP64-residual
    refinement restaged from the FP32 master rows): per-sweep test RMSE
    within 2e-4 of the exact FP64 oracle.
  CG-6: per half-sweep within 1e-4 without the early exit; trajectories
    within twice the distance the oracle's own FP32 execution of the
    recurrence has from its FP64 one.
"""
import numpy as np
import pytest

import cuda_als_amd as ca
from cuda_als_amd import engine as E
from oracle import als_oracle as O

pytestmark = pytest.mark.gpu


def run_gpu(d, f, solve, prec, sweeps, refine=0, lam=0.05, graph=False, batch_work=0, warm=False):
    cfg = E.make_config(f, lam, sweeps, solve, prec, refine=refine, use_graph=graph, batch_work=batch_work,
                        refine_warm=warm)
    with E.Engine(cfg, d) as e:
        e.init_factors()
        r = e.train(sweeps)
        U, V, info = e.get_factors(0), e.get_factors(1), e.info()
    return U, V, r, info


def run_oracle(d, f, solver, sweeps, prec=O.PREC_EXACT, lam=0.05):
    U0, V0 = O.reference_init(d.m, d.n, f)
    cfg = O.config(f, lam, solver, gram_prec=prec, store_fp32=True)
    U, V, r, bad = O.train(cfg, d, sweeps, U0, V0)
    assert bad == 0
    return U, V, r


def test_reference_init_on_device(toy):
    f = 16
    cfg = E.make_config(f)
    with E.Engine(cfg, toy) as e:
        e.init_factors()
        U, V = e.get_factors(0), e.get_factors(1)
    U0, V0 = O.reference_init(toy.m, toy.n, f)
    assert np.array_equal(V, V0.astype(np.float32))
    assert not U.any()


def lapack_fp32_train(d, f, sweeps, lam=0.05):
    """Reference point for FP32 drift: the same ALS with each normal
    equation formed and solved in float32 by numpy/LAPACK (sgemm Gram,
    spotrf/spotrs), factors stored as float32."""
    import scipy.linalg as sl
    U0, V0 = O.reference_init(d.m, d.n, f)
    U, V = U0.astype(np.float32), V0.astype(np.float32)

    def half(ptr, idx, val, Y, X):
        out = np.zeros_like(X)
        for u in range(len(ptr) - 1):
            s, e = ptr[u], ptr[u + 1]
            if e == s:
                continue
            Yu = Y[idx[s:e]]
            A = Yu.T @ Yu + np.float32(lam * (e - s)) * np.eye(f, dtype=np.float32)
            out[u] = sl.cho_solve(sl.cho_factor(A), Yu.T @ val[s:e])
        return out

    rm = []
    for _ in range(sweeps):
        U = half(d.csr_ptr, d.csr_idx, d.csr_val, V, U)
        V = half(d.csc_ptr, d.csc_idx, d.csc_val, U, V)
        rm.append(O.rmse(d.t_row, d.t_col, d.t_val, U, V))
    return U, V, np.array(rm)


@pytest.mark.parametrize("solve", [E.SolveType.CHOL, E.SolveType.LU])
@pytest.mark.parametrize("refine", [0, 1])
def test_fp32_cholesky_matches_oracle(toy, solve, refine):
    """Per-sweep test RMSE within 1e-4 of the FP64 oracle over 5 sweeps, and
    factors within 1e-4 after the first sweep. After 5 sweeps: with one
    FP64-residual refinement step (restaged from the ratings) the factors
    are within 1e-4 too. The plain FP32 solve drifts further on this
    workload's ill-conditioned few-rating items (each half-sweep adds FP32
    Gram accumulation and factorisation error, kappa * eps); its drift is
    bounded by 1.5x that of the same ALS done in float32 with numpy/LAPACK
    (sgemm + spotrf), printed beside it."""
    U, V, r, info = run_gpu(toy, 16, solve, E.Prec.FP32, 5, refine=refine)
    Uo, Vo, ro = run_oracle(toy, 16, O.SOLVE_CHOL, 5)
    drift = max(np.abs(U - Uo).max(), np.abs(V - Vo).max())
    print("gpu", r, "\noracle", ro, "\nfactor drift", drift)
    assert info.failures == 0
    assert np.max(np.abs(r - ro)) <= 1e-4
    if refine:
        assert drift <= 1e-4
    else:
        Ul, Vl, rl = lapack_fp32_train(toy, 16, 5)
        lapack_drift = max(np.abs(Ul - Uo).max(), np.abs(Vl - Vo).max())
        print("float32 LAPACK drift", lapack_drift, "rmse diff", np.abs(rl - ro).max())
        assert drift <= max(1e-4, 1.5 * lapack_drift)
    U1, V1, _, _ = run_gpu(toy, 16, solve, E.Prec.FP32, 1, refine=refine)
    Uo1, Vo1, _ = run_oracle(toy, 16, O.SOLVE_CHOL, 1)
    assert np.max(np.abs(U1 - Uo1)) <= 1e-4 and np.max(np.abs(V1 - Vo1)) <= 1e-4


@pytest.mark.parametrize("f", [16, 100])
@pytest.mark.parametrize("refine", [0, 1])
def test_fp32_half_sweep_error_from_trained_state(toy, f, refine):
    """From the oracle's state after 3 sweeps, one FP32 user half-sweep stays
    within 1e-5 of the FP64 solve on the same (FP32) inputs."""
    U0, V0 = O.reference_init(toy.m, toy.n, f)
    U, V, _, _ = O.train(O.config(f, 0.05, O.SOLVE_CHOL, store_fp32=True), toy, 3, U0, V0)
    with E.Engine(E.make_config(f, 0.05, 1, E.SolveType.CHOL, E.Prec.FP32, refine=refine), toy) as e:
        e.set_factors(0, U.astype(np.float32))
        e.set_factors(1, V.astype(np.float32))
        e.update(ca.ALS_SIDE_USERS)
        Ug = e.get_factors(0)
    Uo, _ = O.update(O.config(f, 0.05, O.SOLVE_CHOL), toy.csr_ptr, toy.csr_idx, toy.csr_val, V, U)
    err = np.abs(Ug - Uo).max()
    print("max err", err)
    assert err <= 1e-5


def run_gpu_restaged(d, f, prec, sweeps, refine):
    """Classic refinement with every row's residual restaged from the FP32
    master rows (the residual of the unrounded system)."""
    cfg = E.make_config(f, 0.05, sweeps, E.SolveType.CHOL, prec, refine=refine)
    with E.Engine(cfg, d) as e:
        for side in (0, 1):
            e.set_schedule(side, restage_max=1 << 30)
        e.init_factors()
        r = e.train(sweeps)
        return e.get_factors(0), e.get_factors(1), r, e.info()


@pytest.mark.parametrize("f", [16, 64, 128])
@pytest.mark.parametrize("prec,oprec", [(E.Prec.FP16, O.PREC_FP16), (E.Prec.BF16, O.PREC_BF16)])
def test_mixed_trajectory_matches_exact_oracle(toy, f, prec, oprec):
    """16-bit Gram (FP16 or BF16), FP32 Cholesky, two refinement steps whose
    FP64 residual is formed from the FP32 master rows, at f = 16, 64 and 128
    on the 1,000 x 500 toy workload: the correction solves on the 16-bit
    factor converge to the unrounded system, so the test RMSE of every sweep
    is within the mixed-precision bar (2e-4) of the exact FP64 oracle, also
    for BF16, whose rounded system is itself further away (the FP64 oracle
    fed BF16-rounded factors, printed, moves by up to ~2e-4 here). The
    factors' largest absolute difference from the oracle's after the five
    sweeps is printed."""
    U, V, r, info = run_gpu_restaged(toy, f, prec, 5, refine=2)
    Uo, Vo, ro = run_oracle(toy, f, O.SOLVE_CHOL, 5)
    _, _, re = run_oracle(toy, f, O.SOLVE_CHOL, 5, prec=oprec)
    print("gpu", r, "\nexact", ro, "\n16-bit-rounded oracle", re, "\ngpu - exact", np.abs(r - ro).max(),
          "rounded - exact", np.abs(re - ro).max(), "factors", np.abs(U - Uo).max(), np.abs(V - Vo).max())
    assert info.failures == 0
    assert np.max(np.abs(r - ro)) <= 2e-4


@pytest.mark.parametrize("f", [16, 64, 128])
@pytest.mark.parametrize("prec,oprec", [(E.Prec.FP16, O.PREC_FP16), (E.Prec.BF16, O.PREC_BF16)])
def test_warm_started_trajectory_matches_oracle(toy, f, prec, oprec):
    """The measured solver (16-bit Gram, FP32 Cholesky, one refinement step
    from the previous iterate, its FP64 residual formed from the Gram tiles,
    light rows through the dual system) at f = 16, 64 and 128 on the toy
    workload, where every user and most items take the dual system: every
    sweep's test RMSE is within the mixed-precision bar (2e-4) of the exact
    FP64 oracle (measured: FP16 up to 2.8e-5, BF16 up to 1.5e-4; its
    distance from the oracle of the 16-bit-formed system is printed)."""
    cfg = E.make_config(f, 0.05, 5, E.SolveType.CHOL, prec, refine=1, refine_warm=True)
    with E.Engine(cfg, toy) as e:
        e.init_factors()
        r = e.train(5)
        info = e.info()
    _, _, ro = run_oracle(toy, f, O.SOLVE_CHOL, 5)
    _, _, re = run_oracle(toy, f, O.SOLVE_CHOL, 5, prec=oprec)
    print("gpu", r, "\nexact", ro, "\n16-bit-rounded oracle", re, "\ngpu - rounded", np.abs(r - re).max(),
          "gpu - exact", np.abs(r - ro).max())
    assert info.failures == 0
    assert np.max(np.abs(r - ro)) <= 2e-4


def test_bf16_half_sweep(toy):
    """One users half-sweep from the reference init with a BF16 Gram. Plain
    FP32 Cholesky: the FP64 solve of the BF16-formed system to FP32 accuracy.
    Two refinement steps restaged from the FP32 master rows: the exact
    system's solve (printed: how far the BF16-formed system's solve is)."""
    f = 16
    U0, V0 = O.reference_init(toy.m, toy.n, f)
    out = {}
    for refine in (0, 2):
        with E.Engine(E.make_config(f, 0.05, 1, E.SolveType.CHOL, E.Prec.BF16, refine=refine), toy) as e:
            e.set_schedule(0, restage_max=1 << 30)
            e.init_factors()
            e.update(ca.ALS_SIDE_USERS)
            out[refine] = e.get_factors(0)
    Ub, _ = O.update(O.config(f, 0.05, O.SOLVE_CHOL, gram_prec=O.PREC_BF16), toy.csr_ptr, toy.csr_idx,
                     toy.csr_val, V0, U0)
    Ux, _ = O.update(O.config(f, 0.05, O.SOLVE_CHOL), toy.csr_ptr, toy.csr_idx, toy.csr_val, V0, U0)
    scale = max(1.0, np.abs(Ux).max())
    print("plain vs bf16 system", np.abs(out[0] - Ub).max(), "refined vs exact", np.abs(out[2] - Ux).max(),
          "bf16 system vs exact", np.abs(Ub - Ux).max())
    assert np.max(np.abs(out[0] - Ub)) <= 1e-5 * scale
    assert np.max(np.abs(out[2] - Ux)) <= 1e-5 * scale


def _cg_half_sweep(d, f, solve, osolve, tol, iters=6):
    U0, V0 = O.reference_init(d.m, d.n, f)
    rng = np.random.default_rng(3)
    U0 = (rng.random((d.m, f)) * 0.1).astype(np.float32).astype(np.float64)   # non-zero warm start
    with E.Engine(E.make_config(f, 0.05, 1, solve, E.Prec.FP32, cg_iters=iters, cg_tol=tol), d) as e:
        e.init_factors()
        e.set_factors(0, U0)
        e.update(ca.ALS_SIDE_USERS)
        U = e.get_factors(0)
    Uo, _ = O.update(O.config(f, 0.05, osolve, cg_iters=iters, cg_tol=tol), d.csr_ptr, d.csr_idx, d.csr_val, V0, U0)
    return U, Uo


@pytest.mark.parametrize("solve,osolve", [(E.SolveType.CUMF_ALS_CG_FP32, O.SOLVE_CG),
                                          (E.SolveType.CUMF_ALS_CG_FP16, O.SOLVE_CG_FP16A)])
def test_cg_half_sweep_matches_oracle_without_early_exit(toy, solve, osolve):
    """With the r.r < tol exit disabled the CG iterate is a smooth function of
    the inputs: FP32 CG must follow the FP64 recurrence (warm start, 6 steps)."""
    U, Uo = _cg_half_sweep(toy, 16, solve, osolve, 0.0)
    err = np.max(np.abs(U - Uo)) / max(1.0, np.abs(Uo).max())
    print("max rel err", err)
    assert err <= 1e-4


def test_cg_early_exit_flips_are_isolated(toy):
    """With tol = 1e-4 a row whose r.r lands near the threshold can take one
    iteration fewer or more in FP32 than in FP64. Those rows are found from
    the FP64 oracle alone: the ones whose result changes when tol moves by
    +-5% (their exit decision depends on r.r within 5% of tol). Every row
    where the GPU differs from the FP64 recurrence by more than 1e-4
    (relative to max(1, |x|)) must be one of them, and they are few."""
    U, Uo = _cg_half_sweep(toy, 16, E.SolveType.CUMF_ALS_CG_FP32, O.SOLVE_CG, 1e-4)
    _, lo = _cg_half_sweep(toy, 16, E.SolveType.CUMF_ALS_CG_FP32, O.SOLVE_CG, 0.95e-4)
    _, hi = _cg_half_sweep(toy, 16, E.SolveType.CUMF_ALS_CG_FP32, O.SOLVE_CG, 1.05e-4)
    near = (np.abs(lo - Uo).max(axis=1) > 0) | (np.abs(hi - Uo).max(axis=1) > 0)
    diff = np.abs(U - Uo).max(axis=1) / np.maximum(1.0, np.abs(Uo).max(axis=1))
    off = diff > 1e-4
    print("rows near the threshold:", int(near.sum()), "rows off by > 1e-4:", int(off.sum()),
          "largest off-threshold difference:", float(diff[~near].max()))
    assert not np.any(off & ~near)
    assert near.mean() <= 0.05


def test_cg_fp32_trajectory(toy):
    """Five sweeps of CG-6 in FP32 against the FP64 recurrence. The truncated
    CG iterate is sensitive to rounding: the oracle's own FP32 execution of
    the same recurrence (ORC_SOLVE_CG32, sums in index order) is ~1e-3 away
    from the FP64 one in test RMSE on this workload. The GPU's distance is
    bounded by twice that spread (and 1e-4 when the spread is smaller)."""
    U, V, r, info = run_gpu(toy, 16, E.SolveType.CUMF_ALS_CG_FP32, E.Prec.FP32, 5)
    _, _, ro = run_oracle(toy, 16, O.SOLVE_CG, 5)
    _, _, r32 = run_oracle(toy, 16, O.SOLVE_CG32, 5)
    spread = np.abs(r32 - ro).max()
    print("gpu", r, "\noracle fp64", ro, "\noracle fp32", r32, "\ngpu - fp64", np.abs(r - ro).max(),
          "fp32 oracle - fp64", spread)
    assert info.failures == 0
    assert np.max(np.abs(r - ro)) <= max(1e-4, 2.0 * spread)


@pytest.mark.parametrize("f", [1, 7, 15, 16, 31, 48, 64, 100, 128])
@pytest.mark.parametrize("prec", [E.Prec.FP32, E.Prec.FP16])
def test_half_sweep_every_tile_count(toy, f, prec):
    """One user half-sweep from the reference init, every 16-block count.
    FP16 Gram with one classic refinement step, whose residual is restaged
    from the FP32 master rows: it lands on the exact system's solution."""
    U0, V0 = O.reference_init(toy.m, toy.n, f)
    cfg = E.make_config(f, 0.05, 1, E.SolveType.CHOL, prec, refine=1 if prec != E.Prec.FP32 else 0)
    with E.Engine(cfg, toy) as e:
        e.init_factors()
        e.update(ca.ALS_SIDE_USERS)
        U = e.get_factors(0)
        assert e.info().failures == 0
    Uo, bad = O.update(O.config(f, 0.05, O.SOLVE_CHOL), toy.csr_ptr, toy.csr_idx, toy.csr_val,
                       V0, U0)
    err = np.max(np.abs(U - Uo)) / max(1.0, np.max(np.abs(Uo)))
    print(f, prec, err)
    assert err <= 1e-4


def test_rmse_kernel(toy):
    rng = np.random.default_rng(1)
    f = 100
    U = rng.standard_normal((toy.m, f)).astype(np.float32) * 0.2
    V = rng.standard_normal((toy.n, f)).astype(np.float32) * 0.2
    with E.Engine(E.make_config(f), toy) as e:
        e.set_factors(0, U)
        e.set_factors(1, V)
        rt, rr = e.rmse(ca.ALS_RMSE_TEST), e.rmse(ca.ALS_RMSE_TRAIN)
    assert abs(rt - O.rmse(toy.t_row, toy.t_col, toy.t_val, U, V)) <= 1e-6 * rt
    rows = toy.train_rows()
    assert abs(rr - O.rmse(rows, toy.csr_idx, toy.csr_val, U, V)) <= 1e-6 * rr


def test_empty_rows_get_zero_factors():
    # users 3 and 7 have no ratings, item 5 has none either
    rng = np.random.default_rng(7)
    m, n = 12, 9
    dense = (rng.random((m, n)) < 0.5) * rng.integers(1, 6, (m, n))
    dense[[3, 7], :] = 0
    dense[:, 5] = 0
    rows, cols = np.nonzero(dense)
    ptr = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=m))])
    d = ca.from_arrays(m, n, ptr, cols, dense[rows, cols].astype(np.float32), rows[:3], cols[:3],
                       dense[rows[:3], cols[:3]].astype(np.float32))
    U, V, r, info = run_gpu(d, 4, E.SolveType.CHOL, E.Prec.FP32, 3)
    assert info.failures == 0
    assert np.isfinite(U).all() and np.isfinite(V).all()
    assert not U[[3, 7]].any() and not V[5].any()
    # The item systems here have 4-6 ratings and condition numbers up to
    # ~290: the FP32 Gram's own rounding (~1e-5 per half-sweep after the
    # solve) compounds over sweeps, so parity is checked one half-sweep at a
    # time from the oracle's state after 2 sweeps, both sides.
    Uo, Vo, ro = run_oracle(d, 4, O.SOLVE_CHOL, 2)
    for side, (p, i, v), Y, X in ((ca.ALS_SIDE_USERS, (d.csr_ptr, d.csr_idx, d.csr_val), Vo, Uo),
                                  (ca.ALS_SIDE_ITEMS, (d.csc_ptr, d.csc_idx, d.csc_val), Uo, Vo)):
        with E.Engine(E.make_config(4, 0.05, 1, E.SolveType.CHOL, E.Prec.FP32), d) as e:
            e.set_factors(0, Uo.astype(np.float32))
            e.set_factors(1, Vo.astype(np.float32))
            e.update(side)
            Xg = e.get_factors(side)
        Xn, _ = O.update(O.config(4, 0.05, O.SOLVE_CHOL), p, i, v, Y.astype(np.float32), X)
        assert np.max(np.abs(Xg - Xn)) <= 1e-4 * max(1.0, np.abs(Xn).max())


def test_16bit_gram_refuses_ratings_beyond_fp16_range():
    """The 16-bit-Gram paths factor with FP16 products, so a row whose
    ratings would put |L^-1 b| (at most sqrt(sum r^2), or sqrt(sum r^2 /
    (lambda n)) below lambda n = 1) outside FP16's range is refused at
    als_set_train; the FP32 path takes the same data."""
    rng = np.random.default_rng(3)
    m, n = 40, 30
    dense = (rng.random((m, n)) < 0.3) * rng.integers(1, 6, (m, n)).astype(np.float32)
    dense[0, :10] = 20000.0  # sum of squares 4e9 on row 0
    rows, cols = np.nonzero(dense)
    ptr = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=m))])
    d = ca.from_arrays(m, n, ptr, cols, dense[rows, cols], rows[:3], cols[:3], dense[rows[:3], cols[:3]])
    for prec in (E.Prec.FP16, E.Prec.BF16):
        with pytest.raises(ca.AlsError, match="16-bit Gram"):
            E.Engine(E.make_config(16, 0.05, 1, E.SolveType.CHOL, prec), d)
    with E.Engine(E.make_config(16, 0.05, 1, E.SolveType.CHOL, E.Prec.FP32), d) as e:
        e.init_factors()
        e.sweep()
        assert np.isfinite(e.get_factors(0)).all()
    # 1000-star ratings stay inside the range
    dense[0, :10] = 1000.0
    rows, cols = np.nonzero(dense)
    d = ca.from_arrays(m, n, ptr, cols, dense[rows, cols], rows[:3], cols[:3], dense[rows[:3], cols[:3]])
    with E.Engine(E.make_config(16, 0.05, 1, E.SolveType.CHOL, E.Prec.FP16), d) as e:
        e.init_factors()
        e.sweep()
        assert e.info().failures == 0 and np.isfinite(e.get_factors(0)).all()


@pytest.mark.parametrize("f,warm", [(16, False), (100, True)])
def test_graph_replay_matches_direct(toy, f, warm):
    """Each sweep replayed from the captured graph (with the RMSE and host
    synchronisation between replays) equals the direct launches; at f = 100
    the graph holds the dual kernel's launches too."""
    U1, V1, r1, _ = run_gpu(toy, f, E.SolveType.CHOL, E.Prec.FP16, 4, refine=1, warm=warm)
    U2, V2, r2, _ = run_gpu(toy, f, E.SolveType.CHOL, E.Prec.FP16, 4, refine=1, graph=True, warm=warm)
    assert np.array_equal(r1, r2)
    assert np.array_equal(U1, U2) and np.array_equal(V1, V2)


def test_dual_kernel_beside_equals_after(monkeypatch):
    """The dual kernel runs beside the update kernel on a second stream; with
    ALS_DUAL_SEQUENTIAL=1 at engine creation it runs after it. The two write
    disjoint rows, so the factors and the RMSE trajectory are identical, in
    direct launches and in a captured graph. netflix-shaped matrix (scaled),
    f = 100, FP16 Gram, one warm-started step: both kernels have work."""
    d = ca.scaled(ca.netflix_shape(), 0.02).generate()
    out = []
    for seq, graph in (("1", False), ("0", False), ("0", True)):
        monkeypatch.setenv("ALS_DUAL_SEQUENTIAL", seq)
        U, V, r, info = run_gpu(d, 100, E.SolveType.CHOL, E.Prec.FP16, 3, refine=1, warm=True, graph=graph)
        assert info.failures == 0 and info.dual_rows[0] > 0 and info.nunits[0] > info.dual_rows[0]
        out.append((U, V, r))
    for U, V, r in out[1:]:
        assert np.array_equal(U, out[0][0]) and np.array_equal(V, out[0][1]) and np.array_equal(r, out[0][2])


def test_schedule_invariance(toy):
    """Batch size changes which wave takes which row, never the result."""
    _, _, r1, i1 = run_gpu(toy, 32, E.SolveType.CHOL, E.Prec.FP32, 2, batch_work=1)
    _, _, r2, i2 = run_gpu(toy, 32, E.SolveType.CHOL, E.Prec.FP32, 2, batch_work=100000)
    assert i1.nbatches[0] != i2.nbatches[0]
    assert np.array_equal(r1, r2)


def test_grid_invariance():
    """The update kernel's grid (persistent waves, set per side) changes which
    wave takes which row, never the result: bit-identical factors at the
    occupancy limit, three quarters of it and an odd small grid, with split
    rows (split_len 300) and light rows through the dual system. Grids
    outside 1 .. the limit are refused."""
    d = ca.scaled(ca.netflix_shape(), 0.01).generate()
    out = []
    for frac in (None, 0.75, "odd"):
        cfg = E.make_config(100, 0.05, 2, E.SolveType.CHOL, E.Prec.FP16, refine=1, refine_warm=True, split_len=300)
        with E.Engine(cfg, d) as e:
            gmax = e.info().grid[0]
            if frac is not None:
                g = int(gmax * frac) if frac != "odd" else 37
                for side in (0, 1):
                    e.set_grid(side, g)
                assert list(e.info().grid) == [g, g]
            with pytest.raises(ca.AlsError):
                e.set_grid(0, gmax + 1)
            with pytest.raises(ca.AlsError):
                e.set_grid(1, 0)
            e.init_factors()
            e.train(2, with_rmse=False)
            info = e.info()
            assert info.failures == 0 and info.nsplit_rows[1] > 0 and info.dual_rows[0] > 0
            out.append((e.get_factors(0), e.get_factors(1)))
    for U, V in out[1:]:
        assert np.array_equal(U, out[0][0]) and np.array_equal(V, out[0][1])


def _run_split(d, f, prec, split_len, sweeps=2):
    cfg = E.make_config(f, 0.05, sweeps, E.SolveType.CHOL, prec, refine=1 if prec != E.Prec.FP32 else 0,
                        split_len=split_len)
    with E.Engine(cfg, d) as e:
        e.init_factors()
        r = e.train(sweeps)
        return e.get_factors(0), e.get_factors(1), r, e.info()


def test_split_rows_match_whole_rows_fp32(toy):
    """Rows cut into parts (partial Gram tiles summed by the last part) give
    the whole-row result up to FP32 summation order."""
    U1, V1, r1, i1 = _run_split(toy, 32, E.Prec.FP32, 8)
    U2, V2, r2, i2 = _run_split(toy, 32, E.Prec.FP32, 1 << 30)
    print("split rows", i1.nsplit_rows[0], i1.nsplit_rows[1], "units", i1.nunits[0], i1.nunits[1])
    assert i1.nsplit_rows[1] > 100 and i2.nsplit_rows[1] == 0
    assert i1.failures == 0
    assert np.max(np.abs(U1 - U2)) <= 1e-4 and np.max(np.abs(V1 - V2)) <= 1e-4
    assert np.max(np.abs(r1 - r2)) <= 1e-5


@pytest.mark.parametrize("side", [ca.ALS_SIDE_USERS, ca.ALS_SIDE_ITEMS])
def test_split_rows_match_whole_rows_fp16_half_sweep(toy, side):
    """FP16 Gram, identical 16-bit inputs: split and whole rows differ only by
    the FP32 order of the partial sums. (Across sweeps that order difference
    can flip an FP16 rounding of a factor, so trajectories are not compared.)"""
    rng = np.random.default_rng(5)
    f = 32
    U0 = (rng.random((toy.m, f)) * 0.3).astype(np.float32)
    V0 = (rng.random((toy.n, f)) * 0.3).astype(np.float32)
    out = []
    for split_len in (8, 1 << 30):
        cfg = E.make_config(f, 0.05, 1, E.SolveType.CHOL, E.Prec.FP16, refine=1, split_len=split_len)
        with E.Engine(cfg, toy) as e:
            # every row refines against its FP32-accumulated tiles (split rows
            # always do), so the two differ only in the order of the sums
            e.set_schedule(side, restage_max=0)
            e.set_factors(0, U0)
            e.set_factors(1, V0)
            e.update(side)
            out.append(e.get_factors(side))
            assert e.info().failures == 0
    scale = max(1.0, np.abs(out[1]).max())
    assert np.max(np.abs(out[0] - out[1])) <= 1e-5 * scale


def test_split_rows_are_deterministic(toy):
    """Which part arrives last varies from run to run; the partials are summed
    in part order, so the result does not."""
    U1, V1, r1, _ = _run_split(toy, 32, E.Prec.FP16, 5, sweeps=3)
    U2, V2, r2, _ = _run_split(toy, 32, E.Prec.FP16, 5, sweeps=3)
    assert np.array_equal(U1, U2) and np.array_equal(V1, V2) and np.array_equal(r1, r2)


def test_autotune_keeps_state_and_results():
    """The autotuner times candidate schedules (split length x batch size,
    then the grid) on the engine's own data: the
    factors it leaves behind are the ones it found, the chosen schedule is
    the fastest candidate (the default is one of them), and a sweep on the
    tuned schedule matches a sweep on the default one."""
    d = ca.scaled(ca.netflix_shape(), 0.04).generate()
    f = 32
    cfg = E.make_config(f, 0.05, 1, E.SolveType.CHOL, E.Prec.FP16, refine=1)
    with E.Engine(cfg, d) as e0, E.Engine(cfg, d) as e1:
        e0.init_factors()
        e1.init_factors()
        e1.sweep()
        U_before, V_before = e1.get_factors(0), e1.get_factors(1)
        r = e1.autotune()
        print(r)
        assert np.array_equal(e1.get_factors(0), U_before) and np.array_equal(e1.get_factors(1), V_before)
        for sd in (0, 1):
            assert r["candidates"][sd] >= 2
            assert r["best_ms"][sd] <= r["default_ms"][sd]
            assert r["split_len"][sd] >= 256 and r["batch_work"][sd] >= 1
        assert list(e1.info().split_len) == r["split_len"]
        gmax = e0.info().grid[0]
        assert list(e1.info().grid) == r["grid"] and all(gmax <= 2 * g <= 2 * gmax for g in r["grid"])
        # The tuned restage threshold moves rows between the two refinement
        # residuals, which converge to different systems (restaged from the
        # FP32 rows: the exact one; the FP32-accumulated 16-bit tiles: the
        # FP16-rounded one, ~5e-4 away on the lightest rows here). For the
        # comparison both engines keep the default threshold, 32 (T + 1).
        for sd in (0, 1):
            e1.set_schedule(sd, restage_max=32 * (e1.info().T + 1))
        e0.sweep()
        e0.sweep()
        e1.sweep()
        U0, V0, U1, V1 = e0.get_factors(0), e0.get_factors(1), e1.get_factors(0), e1.get_factors(1)
        assert e1.info().failures == 0
        # a different split length cuts different parts: FP32 summation order
        print("tuned vs default", np.abs(U0 - U1).max(), np.abs(V0 - V1).max())
        assert np.max(np.abs(U0 - U1)) <= 2e-5 * max(1.0, np.abs(U0).max())
        assert np.max(np.abs(V0 - V1)) <= 2e-5 * max(1.0, np.abs(V0).max())


def test_autotune_warm_tries_split_lengths_under_each_variant():
    """With the warm-started solve at 7 tiles (f = 100) both update kernel
    variants are built, and the autotuner times the split length x batch
    size candidates under each: it leaves the factors as it found them,
    picks a built variant, its best time is no worse than the default
    schedule's, and a sweep on the tuned schedule matches one on the
    default schedule (a different split length only reorders FP32 sums)."""
    d = ca.scaled(ca.netflix_shape(), 0.05).generate()
    cfg = E.make_config(100, 0.05, 1, E.SolveType.CHOL, E.Prec.FP16, refine=1, refine_warm=True)
    with E.Engine(cfg, d) as e0, E.Engine(cfg, d) as e1:
        assert e1.info().grid_variant1 > 0
        e0.init_factors()
        e1.init_factors()
        U_before, V_before = e1.get_factors(0), e1.get_factors(1)
        r = e1.autotune()
        print(r)
        assert np.array_equal(e1.get_factors(0), U_before) and np.array_equal(e1.get_factors(1), V_before)
        assert e1.info().failures == 0
        for sd in (0, 1):
            assert r["kernel_variant"][sd] in (0, 1)
            assert r["best_ms"][sd] <= r["default_ms"][sd]
        assert list(e1.info().kernel_variant) == r["kernel_variant"]
        assert list(e1.info().split_len) == r["split_len"]
        # The tuned dual limit moves rows between the dual system (a plain
        # FP32 solve) and the primal warm-started step, which from the
        # random initial factors differ by up to ~1e-4 on the lightest,
        # ill-conditioned items. For the comparison both engines keep the
        # default limit.
        for sd in (0, 1):
            e1.set_dual_limit(sd, e0.info().dual_limit[sd])
        e0.sweep()
        e1.sweep()
        U0, V0, U1, V1 = e0.get_factors(0), e0.get_factors(1), e1.get_factors(0), e1.get_factors(1)
        assert e1.info().failures == 0
        print("tuned vs default", np.abs(U0 - U1).max(), np.abs(V0 - V1).max())
        assert np.max(np.abs(U0 - U1)) <= 2e-5 * max(1.0, np.abs(U0).max())
        assert np.max(np.abs(V0 - V1)) <= 2e-5 * max(1.0, np.abs(V0).max())


@pytest.mark.parametrize("prec,refine,warm", [(E.Prec.BF16, 0, False), (E.Prec.FP16, 0, False),
                                              (E.Prec.FP16, 1, True)])
def test_group_single_launch_equals_separate_runs(prec, refine, warm):
    """Three models, f = 32 / 64 / 128, on one ml10M-shaped matrix (scaled),
    updated by ONE graph launch per half-sweep (each model's kernels on its
    own stream): every model's factors equal its own separate run bit for
    bit."""
    d = ca.scaled(ca.ml10m_shape(), 0.1).generate()
    fs = (32, 64, 128)
    cfgs = [E.make_config(f, 0.05, 1, E.SolveType.CHOL, prec, refine=refine, refine_warm=warm) for f in fs]
    sep = [E.Engine(c, d) for c in cfgs]
    grp = [E.Engine(c, d) for c in cfgs]
    for e in sep + grp:
        e.init_factors()
    with E.Group(grp) as g:
        for _ in range(2):
            g.sweep()
            for e in sep:
                e.sweep()
    for es, eg, f in zip(sep, grp, fs):
        for side in (0, 1):
            a, b = es.get_factors(side), eg.get_factors(side)
            assert np.array_equal(a, b), (f, side, np.abs(a - b).max())
        assert eg.info().failures == 0
        assert abs(es.rmse() - eg.rmse()) == 0.0
    for e in sep + grp:
        e.close()


def test_group_mixed_models_and_rejections(toy):
    """Models of any f, precision and solver group together (here FP16 f=32,
    BF16 f=64, f=8, and CG) and each equals its own run; engines over
    another rating matrix, a repeated engine or more than four are refused."""
    cfgs = [E.make_config(32, prec=E.Prec.FP16), E.make_config(64, prec=E.Prec.BF16),
            E.make_config(8, prec=E.Prec.FP16), E.make_config(48, solve=E.SolveType.CUMF_ALS_CG_FP32)]
    sep = [E.Engine(c, toy) for c in cfgs]
    grp = [E.Engine(c, toy) for c in cfgs]
    for e in sep + grp:
        e.init_factors()
    with E.Group(grp) as g:
        g.sweep()
    for e in sep:
        e.sweep()
    for es, eg in zip(sep, grp):
        assert np.array_equal(es.get_factors(0), eg.get_factors(0))
        assert np.array_equal(es.get_factors(1), eg.get_factors(1))
    other = ca.scaled(ca.ml10m_shape(), 0.02).generate()
    eo = E.Engine(E.make_config(32, prec=E.Prec.FP16), other)
    with pytest.raises(ca.AlsError):
        E.Group([grp[0], eo])          # another rating matrix
    with pytest.raises(ca.AlsError):
        E.Group([grp[0], grp[0]])      # the same engine twice
    with pytest.raises(ca.AlsError):
        E.Group(grp + [eo])            # five engines
    for e in sep + grp + [eo]:
        e.close()


def test_warm_started_refinement(toy):
    """refine_warm: the first refinement step starts from the row's previous
    value x0, with the FP64 residual b - A x0 formed before the
    factorisation from the 16-bit Gram. Its error is kappa * eps * |x* - x0|:
    from a trained state a half-sweep lands on the FP64 solve of the
    FP16-rounded system; a classic step after it (its residual restaged from
    the FP32 master rows), or a classic step alone, lands on the exact
    system's solve. The trajectory meets the mixed-precision bar."""
    f = 16
    U, V, r, info = None, None, None, None
    with E.Engine(E.make_config(f, 0.05, 5, E.SolveType.CHOL, E.Prec.FP16, refine=1, refine_warm=True), toy) as e:
        e.init_factors()
        r = e.train(5)
        assert e.info().failures == 0
    _, _, ro = run_oracle(toy, f, O.SOLVE_CHOL, 5)
    _, _, re = run_oracle(toy, f, O.SOLVE_CHOL, 5, prec=O.PREC_FP16)
    print("warm", r, "\nexact", ro, "\nfp16-rounded", re)
    assert np.max(np.abs(r - ro)) <= 2e-4 and np.max(np.abs(r - re)) <= 1e-4
    # one half-sweep from the oracle's trained state (x0 = the state itself)
    U0, V0 = O.reference_init(toy.m, toy.n, f)
    Uo, Vo, _, _ = O.train(O.config(f, 0.05, O.SOLVE_CHOL, gram_prec=O.PREC_FP16, store_fp32=True), toy, 3, U0, V0)
    errs = {}
    for warm, steps in ((True, 1), (False, 1), (True, 2)):
        with E.Engine(E.make_config(f, 0.05, 1, E.SolveType.CHOL, E.Prec.FP16, refine=steps, refine_warm=warm),
                      toy) as e:
            e.set_factors(0, Uo.astype(np.float32))
            e.set_factors(1, Vo.astype(np.float32))
            e.update(ca.ALS_SIDE_USERS)
            Ug = e.get_factors(0)
        oprec = O.PREC_FP16 if (warm, steps) == (True, 1) else O.PREC_EXACT
        Un, _ = O.update(O.config(f, 0.05, O.SOLVE_CHOL, gram_prec=oprec), toy.csr_ptr, toy.csr_idx,
                         toy.csr_val, Vo.astype(np.float32), Uo)
        errs[(warm, steps)] = np.abs(Ug - Un).max() / max(1.0, np.abs(Un).max())
    print(errs)
    assert errs[(True, 1)] <= 1e-5 and errs[(False, 1)] <= 1e-5 and errs[(True, 2)] <= 1e-5
    # from x0 = 0 the warm step is the plain FP32 solve of the first sweep
    with E.Engine(E.make_config(f, 0.05, 1, E.SolveType.CHOL, E.Prec.FP16, refine=1, refine_warm=True), toy) as e, \
            E.Engine(E.make_config(f, 0.05, 1, E.SolveType.CHOL, E.Prec.FP16), toy) as p:
        e.init_factors()
        p.init_factors()
        e.update(ca.ALS_SIDE_USERS)
        p.update(ca.ALS_SIDE_USERS)
        a, b = e.get_factors(0), p.get_factors(0)
        assert np.max(np.abs(a - b)) <= 1e-5 * max(1.0, np.abs(b).max())


@pytest.mark.parametrize("f", [64, 100])
@pytest.mark.parametrize("prec,oprec", [(E.Prec.FP16, O.PREC_FP16), (E.Prec.BF16, O.PREC_BF16)])
def test_dual_system_for_light_rows(toy, f, prec, oprec):
    """With a 16-bit Gram and one warm-started refinement step, whole rows
    with fewer than 16 min(T - 2, 5) ratings (48 at f = 64, 80 at f = 100)
    are solved in the space of their ratings, x = Y^T (Y Y^T + mu I)^-1 r,
    by a plain FP32 Cholesky solve (the refinement step is the primal
    rows'). Every user of the toy workload (at most 40 ratings) and most of
    its items take that path. One half-sweep of each side from the oracle's
    trained state matches the FP64 solve of the 16-bit-rounded system within
    1e-5 relative to max(1, |x|); the rows through the dual system land
    within ~3e-7 (bar 1e-6): its matrix is at least lambda n I, so the FP32
    solve needs no refinement there."""
    assert np.diff(toy.csr_ptr).max() < 48
    U0, V0 = O.reference_init(toy.m, toy.n, f)
    Uo, Vo, _, _ = O.train(O.config(f, 0.05, O.SOLVE_CHOL, gram_prec=oprec, store_fp32=True), toy, 3, U0, V0)
    with E.Engine(E.make_config(f, 0.05, 1, E.SolveType.CHOL, prec, refine=1, refine_warm=True), toy) as e:
        e.set_factors(0, Uo.astype(np.float32))
        e.set_factors(1, Vo.astype(np.float32))
        e.update(ca.ALS_SIDE_USERS)
        Ug = e.get_factors(0)
        e.update(ca.ALS_SIDE_ITEMS)
        Vg = e.get_factors(1)
        assert e.info().failures == 0
    cfg = O.config(f, 0.05, O.SOLVE_CHOL, gram_prec=oprec)
    Un, _ = O.update(cfg, toy.csr_ptr, toy.csr_idx, toy.csr_val, Vo.astype(np.float32), Uo)
    Vn, _ = O.update(cfg, toy.csc_ptr, toy.csc_idx, toy.csc_val, Ug.astype(np.float64), Vo)
    eu = np.abs(Ug - Un).max(axis=1) / np.maximum(1.0, np.abs(Un).max(axis=1))
    ev = np.abs(Vg - Vn).max(axis=1) / np.maximum(1.0, np.abs(Vn).max(axis=1))
    light = np.diff(toy.csc_ptr) < 48
    print("users", eu.max(), "light items", ev[light].max(), "other items", ev[~light].max())
    assert eu.max() <= 1e-6 and ev[light].max() <= 1e-6 and ev.max() <= 1e-5


@pytest.mark.parametrize("f", [64, 100])
@pytest.mark.parametrize("prec,refine,warm,oprec,tol", [(E.Prec.FP32, 0, False, O.PREC_EXACT, 1e-4),
                                                        (E.Prec.FP16, 0, False, O.PREC_FP16, 1e-4),
                                                        (E.Prec.FP16, 1, True, O.PREC_FP16, 1e-5)])
def test_dual_system_tile_boundaries(f, prec, refine, warm, oprec, tol):
    """Rows whose rating count puts the augmented row n at every position
    the dual system distinguishes: first and last row of a tile (n = 15, 16,
    31, 32, 47, 48, 63, 64, 79), the one-rating row, the empty row, and rows
    just past the dual limit (48 at f = 64, 80 at f = 100) or far past it,
    which take the primal system. One users half-sweep from random factors
    against the FP64 oracle (relative to max(1, |x|)); plain solves 1e-4;
    with the warm-started step 1e-5 (its primal rows refined, its dual rows
    a plain FP32 solve of a system at least lambda n I)."""
    rng = np.random.default_rng(17)
    counts = np.repeat([0, 1, 2, 14, 15, 16, 17, 30, 31, 32, 33, 46, 47, 48, 49, 63, 64, 65, 79, 80, 81, 100], 3)
    m, n_items = len(counts), 400
    cols = np.concatenate([np.sort(rng.choice(n_items, c, replace=False)) for c in counts]).astype(np.int32)
    ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    vals = rng.integers(1, 6, cols.size).astype(np.float32)
    rows = np.repeat(np.arange(m), counts)
    d = ca.from_arrays(m, n_items, ptr, cols, vals, rows[:5], cols[:5], vals[:5])
    U0 = (0.3 * rng.standard_normal((m, f))).astype(np.float32)
    V0 = (0.3 * rng.standard_normal((n_items, f))).astype(np.float32)
    with E.Engine(E.make_config(f, 0.05, 1, E.SolveType.CHOL, prec, refine=refine, refine_warm=warm), d) as e:
        e.set_factors(0, U0)
        e.set_factors(1, V0)
        e.update(ca.ALS_SIDE_USERS)
        Ug = e.get_factors(0)
        assert e.info().failures == 0
    Un, _ = O.update(O.config(f, 0.05, O.SOLVE_CHOL, gram_prec=oprec), d.csr_ptr, d.csr_idx, d.csr_val, V0,
                     U0.astype(np.float64))
    err = np.abs(Ug - Un).max(axis=1) / np.maximum(1.0, np.abs(Un).max(axis=1))
    print({int(c): float(e) for c, e in zip(counts[::3], err.reshape(-1, 3).max(axis=1))})
    assert not Ug[counts == 0].any()
    assert err.max() <= tol


def test_dual_limit_moves_rows_between_systems():
    """The per-side dual limit decides which rows solve the n x n dual system
    and which the f x f one: at limits 80 (the built maximum at f = 100), 48,
    16 and 0 (none), one warm-started users half-sweep from random factors
    takes exactly the rows below the limit (and the empty ones) through the
    dual kernel, and every row matches the FP64 solve of the FP16-rounded
    system within 1e-5 relative to max(1, |x|). Limits above the maximum
    are refused."""
    rng = np.random.default_rng(29)
    f = 100
    counts = np.repeat([0, 1, 15, 16, 17, 40, 47, 48, 49, 63, 64, 79, 80, 81, 150], 3)
    m, n_items = len(counts), 400
    cols = np.concatenate([np.sort(rng.choice(n_items, c, replace=False)) for c in counts]).astype(np.int32)
    ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    vals = rng.integers(1, 6, cols.size).astype(np.float32)
    rows = np.repeat(np.arange(m), counts)
    d = ca.from_arrays(m, n_items, ptr, cols, vals, rows[:5], cols[:5], vals[:5])
    U0 = (0.3 * rng.standard_normal((m, f))).astype(np.float32)
    V0 = (0.3 * rng.standard_normal((n_items, f))).astype(np.float32)
    Un, _ = O.update(O.config(f, 0.05, O.SOLVE_CHOL, gram_prec=O.PREC_FP16), d.csr_ptr, d.csr_idx, d.csr_val, V0,
                     U0.astype(np.float64))
    cfg = E.make_config(f, 0.05, 1, E.SolveType.CHOL, E.Prec.FP16, refine=1, refine_warm=True)
    with E.Engine(cfg, d) as e:
        info = e.info()
        assert info.dual_max == 80 and list(info.dual_limit) == [80, 80]
        with pytest.raises(ca.AlsError):
            e.set_dual_limit(0, 81)
        for limit in (80, 48, 16, 0):
            e.set_dual_limit(0, limit)
            e.set_factors(0, U0)
            e.set_factors(1, V0)
            e.update(ca.ALS_SIDE_USERS)
            Ug = e.get_factors(0)
            info = e.info()
            assert info.dual_limit[0] == limit and info.failures == 0
            assert info.dual_rows[0] == (int(np.sum(counts < limit)) if limit else 0)
            err = np.abs(Ug - Un).max(axis=1) / np.maximum(1.0, np.abs(Un).max(axis=1))
            print(limit, err.max())
            assert not Ug[counts == 0].any()
            assert err.max() <= 1e-5


@pytest.mark.parametrize("f", [16, 63, 100, 127])
@pytest.mark.parametrize("prec,oprec", [(E.Prec.FP16, O.PREC_FP16), (E.Prec.BF16, O.PREC_BF16)])
@pytest.mark.parametrize("split_len", [0, 40])
def test_gram_chunk_boundaries(f, prec, oprec, split_len):
    """The 16-bit Gram walks a row in 32-rating chunks with the next chunk's
    gathers in flight (f <= 127). Rows end at every chunk position that
    loop distinguishes: one rating, a chunk's last and first slot (31, 32,
    33, ...), one to six chunks, an odd and an even chunk count; f = 63
    and 127 fill all 64 lanes with gather tasks, f = 16 leaves most idle.
    split_len 40 cuts the longer rows into parts of 40 ratings (a full and
    a short chunk per part) summed by the last part. At f = 100 and 127 the
    rows under 80 ratings take the dual system instead. One users
    half-sweep from random factors against the FP64 oracle on the same
    16-bit-rounded inputs, relative to max(1, |x|): 1e-4."""
    rng = np.random.default_rng(23)
    counts = np.repeat([1, 31, 32, 33, 63, 64, 65, 95, 96, 97, 127, 128, 129, 160, 191, 192, 193, 250], 2)
    m, n_items = len(counts), 600
    cols = np.concatenate([np.sort(rng.choice(n_items, c, replace=False)) for c in counts]).astype(np.int32)
    ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    vals = rng.integers(1, 6, cols.size).astype(np.float32)
    rows = np.repeat(np.arange(m), counts)
    d = ca.from_arrays(m, n_items, ptr, cols, vals, rows[:5], cols[:5], vals[:5])
    U0 = (0.3 * rng.standard_normal((m, f))).astype(np.float32)
    V0 = (0.3 * rng.standard_normal((n_items, f))).astype(np.float32)
    cfg = E.make_config(f, 0.05, 1, E.SolveType.CHOL, prec, split_len=split_len or (1 << 30))
    with E.Engine(cfg, d) as e:
        e.set_factors(0, U0)
        e.set_factors(1, V0)
        e.update(ca.ALS_SIDE_USERS)
        Ug = e.get_factors(0)
        info = e.info()
        assert info.failures == 0
    if split_len:
        assert info.nsplit_rows[0] >= 20
    Un, _ = O.update(O.config(f, 0.05, O.SOLVE_CHOL, gram_prec=oprec), d.csr_ptr, d.csr_idx, d.csr_val, V0,
                     U0.astype(np.float64))
    err = np.abs(Ug - Un).max(axis=1) / np.maximum(1.0, np.abs(Un).max(axis=1))
    print(f, prec, split_len, {int(c): float(x) for c, x in zip(counts[::2], err.reshape(-1, 2).max(axis=1))})
    assert err.max() <= 1e-4
