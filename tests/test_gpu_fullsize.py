"""Full-size checks at the netflix-shaped workload (480,189 x 17,770,
99,072,112 train ratings, f = 100): the shape the benchmark runs.

The FP64 oracle cannot redo whole half-sweeps here in test time, so parity is
checked on sampled rows from the GPU's own state (uniform users, and the
heaviest items, which the schedule splits across waves), plus the test RMSE
recomputed on the host from the downloaded factors.
Tolerances (relative to max(1, |x|)): mixed path (FP16 Gram, FP32 Cholesky,
one FP64-residual refinement step, classic or warm-started) 5e-5 against the
FP64 solve of the system the refinement converges to: rows whose residual is
restaged from the FP32 master rows (classic, whole rows up to 256 ratings)
against the exact system; rows refined against their FP32-accumulated 16-bit
tiles (long and split rows, every warm-started row) and the light rows solved
through their dual system (a plain FP32 solve) against the FP16-rounded
system, whose accumulation error they keep (~1e-5 on the heaviest items).
FP32 path 1e-4 against the exact FP64 solve.
"""
from pathlib import Path

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
    classic = refine > 0 and not warm
    restage_max = 256   # classic refinement: whole rows up to this many ratings restage their residual
    with E.Engine(E.make_config(f, 0.05, 1, E.SolveType.CHOL, prec, refine=refine, refine_warm=warm), d) as e:
        if classic:
            for side in (0, 1):
                e.set_schedule(side, restage_max=restage_max)
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
    exact = O.config(f, 0.05, O.SOLVE_CHOL)

    def expected(rows, ptr, idx, val, Y, split_len):
        """The oracle solve each row converges to: classic refinement of a
        whole row up to restage_max ratings restages its residual from the
        FP32 master rows (the exact system); other rows refine against
        their 16-bit-formed tiles (that system)."""
        want = O.update_rows(cfg, rows, ptr, idx, val, Y)
        if classic:
            n = np.diff(ptr)[rows]
            restaged = (n <= restage_max) & (n <= split_len)
            want[restaged] = O.update_rows(exact, rows[restaged], ptr, idx, val, Y)
        return want

    Uo = expected(users, d.csr_ptr, d.csr_idx, d.csr_val, V, info.split_len[0])
    eu = np.abs(U1[users] - Uo).max(axis=1) / np.maximum(1.0, np.abs(Uo).max(axis=1))
    Vo = expected(items, d.csc_ptr, d.csc_idx, d.csc_val, U1.astype(np.float64), info.split_len[1])
    ei = np.abs(V1[items] - Vo).max(axis=1) / np.maximum(1.0, np.abs(Vo).max(axis=1))
    print("users max rel err", eu.max(), "items max rel err", ei.max())
    assert eu.max() <= tol and ei.max() <= tol
    # test RMSE: FP64 recomputation from the factors
    pred = np.einsum("ij,ij->i", U1[d.t_row].astype(np.float64), V1[d.t_col].astype(np.float64))
    ref = np.sqrt(np.mean((d.t_val - pred) ** 2))
    assert abs(rmse - ref) <= 1e-6 * ref


# ---------------------------------------------------------------------------
# Ten sweeps from the reference init against the FP64 oracle's trajectory of
# the same workload (tests/golden/netflix_trajectory.json, written by
# tests/golden/make_netflix_trajectory.py: FP64 arithmetic, factors stored
# as float32 after every half-sweep, exact solve and CG-6).
GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def trajectory(netflix):
    import hashlib
    import json
    fix = json.loads((GOLDEN / "netflix_trajectory.json").read_text())
    h = hashlib.sha256()
    for a in (netflix.csr_ptr, netflix.csr_idx, netflix.csr_val, netflix.csc_ptr, netflix.csc_idx, netflix.csc_val,
              netflix.t_row, netflix.t_col, netflix.t_val):
        h.update(np.ascontiguousarray(a).tobytes())
    assert h.hexdigest() == fix["data_sha256"], "the generated workload is not the one the fixture was made from"
    return fix, np.load(GOLDEN / "netflix_factors.npz")


# name, solver, Gram precision, refinement steps, warm start, oracle trajectory, RMSE bound, factor bound
TRAJ = [
    ("fp32_chol", E.SolveType.CHOL, E.Prec.FP32, 0, False, "chol", 1e-4, 1e-4),
    ("mixed_fp16_warm", E.SolveType.CHOL, E.Prec.FP16, 1, True, "chol", 2e-4, None),
    ("mixed_bf16_warm", E.SolveType.CHOL, E.Prec.BF16, 1, True, "chol", 2e-4, None),
    ("fp32_cg6", E.SolveType.CUMF_ALS_CG_FP32, E.Prec.FP32, 0, False, "cg6", 1e-4, None),
]


@pytest.mark.parametrize("name,solve,prec,refine,warm,oname,rtol,ftol", TRAJ, ids=[t[0] for t in TRAJ])
def test_netflix_trajectory_matches_oracle(netflix, trajectory, name, solve, prec, refine, warm, oname, rtol, ftol):
    """Per-sweep test RMSE of 10 sweeps from the reference init within the
    bound of the path (FP32: 1e-4; mixed precision: 2e-4; FP32 CG-6 against
    the oracle's CG-6: 1e-4). Factor rows of the seeded sample after sweeps
    1 and 10: within 1e-4 for the FP32 exact solve; the distance of the
    other paths is printed."""
    fix, fac = trajectory
    d, f = netflix, fix["f"]
    users, items = np.array(fix["sample_users"]), np.array(fix["sample_items"])
    cfg = E.make_config(f, fix["lambda"], fix["sweeps"], solve, prec, refine=refine, refine_warm=warm)
    got, frows = [], {}
    with E.Engine(cfg, d) as e:
        e.init_factors()
        for s in range(1, fix["sweeps"] + 1):
            e.sweep()
            got.append(e.rmse())
            if s in fix["check_sweeps"]:
                frows[s] = (e.get_factors(0)[users], e.get_factors(1)[items])
        assert e.info().failures == 0
    want = np.array(fix[f"rmse_{oname}"])
    got = np.array(got)
    print(name, "gpu", got.tolist(), "\noracle", want.tolist(), "\nmax |diff|", np.abs(got - want).max())
    for s, (Ug, Vg) in frows.items():
        eu = np.abs(Ug - fac[f"{oname}_U_s{s}"]).max()
        ev = np.abs(Vg - fac[f"{oname}_V_s{s}"]).max()
        print(f"sweep {s}: sampled factor max |diff| users {eu:.3e} items {ev:.3e}")
        if ftol is not None:
            assert eu <= ftol and ev <= ftol, (s, eu, ev)
    assert np.abs(got - want).max() <= rtol
