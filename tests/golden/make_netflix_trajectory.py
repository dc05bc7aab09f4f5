"""Writes the FP64-oracle fixture of the netflix-shaped workload.

Workload: 480,189 users x 17,770 items, 99,072,112 train ratings (1,408,395
held out), seed 42, f = 100, lambda = 0.05, the reference's init rule
(srand(0), V = 0.2 rand()/RAND_MAX, U = 0; als_model.cu:1772-1777), 10
sweeps (config/netflix.ini:18 als_iters = 10). Two solvers, both with FP64
arithmetic and the factors stored as float32 after every half-sweep (the
reference stores FP32 factors):
  chol  exact solve of every normal equation (the reference's LU solve,
        als_model.cu:1570-1644, on the same SPD systems)
  cg6   the cuMF CG recurrence, 6 iterations, exit on r.r < 1e-4, warm start
        (cumf_als_cg/cg.cu:37-232; CG_ITER als_model.cu:16; the reference's
        production solver, config/netflix.ini:37)

Stored (netflix_trajectory.json + netflix_factors.npz): the per-sweep test
RMSE of both solvers, the sha256 of the generated arrays (the GPU tests
regenerate the workload and check it is the same), and the factor rows of a
seeded sample of users and items after sweeps 1 and 10.

    python tests/golden/make_netflix_trajectory.py      # ~10-20 min on 8 cores

Built libraries needed: cuda_als_amd/libals_mi355x.so (its host-side data
generator only; no GPU) and oracle/libals_oracle.so.
"""
from __future__ import annotations

import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

F, LAM, SWEEPS = 100, 0.05, 10
CHECK_SWEEPS = (1, 10)


def data_digest(d) -> str:
    h = hashlib.sha256()
    for a in (d.csr_ptr, d.csr_idx, d.csr_val, d.csc_ptr, d.csc_idx, d.csc_val, d.t_row, d.t_col, d.t_val):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def sample_rows(d):
    """Users: 24 uniform, the 8 heaviest, 8 under 80 ratings (the rows the
    16-bit path solves through their dual system). Items: the 8 heaviest and
    24 uniform."""
    rng = np.random.default_rng(2024)
    cu = np.diff(d.csr_ptr)
    users = np.concatenate([rng.choice(d.m, 24, replace=False), np.argsort(cu, kind="stable")[-8:],
                            rng.choice(np.where((cu > 0) & (cu < 80))[0], 8, replace=False)])
    ci = np.diff(d.csc_ptr)
    items = np.concatenate([np.argsort(ci, kind="stable")[-8:], rng.choice(d.n, 24, replace=False)])
    return np.unique(users).astype(np.int64), np.unique(items).astype(np.int64)


def main() -> int:
    import cuda_als_amd as ca
    from oracle import als_oracle as O

    t0 = time.perf_counter()
    d = ca.netflix_shape().generate()
    digest = data_digest(d)
    print(f"generated in {time.perf_counter() - t0:.1f}s, sha256 {digest}", flush=True)
    users, items = sample_rows(d)
    U0, V0 = O.reference_init(d.m, d.n, F)
    out = {"workload": f"netflix-shape {d.m}x{d.n}, {d.nnz} train / {d.n_test} test ratings, seed 42",
           "f": F, "lambda": LAM, "sweeps": SWEEPS, "data_sha256": digest,
           "init": "reference: srand(0), V = 0.2*rand()/RAND_MAX in memory order, U = 0",
           "oracle": "oracle/als_oracle.c, FP64 arithmetic, factors rounded to float32 after every half-sweep",
           "threads": int(O.lib.orc_num_threads()), "sample_users": users.tolist(), "sample_items": items.tolist(),
           "check_sweeps": list(CHECK_SWEEPS)}
    arrays = {}
    for name, solver in (("chol", O.SOLVE_CHOL), ("cg6", O.SOLVE_CG)):
        cfg = O.config(F, LAM, solver, cg_iters=6, cg_tol=1e-4, store_fp32=True)
        U, V = U0.copy(), V0.copy()
        rmse, secs = [], []
        for s in range(1, SWEEPS + 1):
            t = time.perf_counter()
            U, V, r, bad = O.train(cfg, d, 1, U, V)
            secs.append(time.perf_counter() - t)
            assert bad == 0, (name, s, bad)
            rmse.append(float(r[0]))
            print(f"{name} sweep {s}: test RMSE {r[0]:.9f} ({secs[-1]:.1f}s)", flush=True)
            if s in CHECK_SWEEPS:
                arrays[f"{name}_U_s{s}"] = U[users].astype(np.float32)
                arrays[f"{name}_V_s{s}"] = V[items].astype(np.float32)
                assert np.array_equal(arrays[f"{name}_U_s{s}"], U[users])   # float32 values
        out[f"rmse_{name}"] = rmse
        out[f"seconds_per_sweep_{name}"] = float(np.median(secs))
    np.savez_compressed(HERE / "netflix_factors.npz", **arrays)
    (HERE / "netflix_trajectory.json").write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", HERE / "netflix_trajectory.json", HERE / "netflix_factors.npz")
    return 0


if __name__ == "__main__":
    sys.exit(main())
