"""Differentiable half-sweep: the engine's adjoint (through the C ABI, wrapped
in torch.autograd) against the FP64 oracle gradient and FP64 central finite
differences (h = 1e-6), on a seeded 2,000-user subsample of the ml-latest
shaped workload (283,228 x 58,098, 27.75M ratings, f = 64, lambda_u = 0.05).

Loss: held-out RMSE of the subsampled users, L = sqrt(mean((r - x_u.y_i)^2)),
which depends on lambda through x_u and on Y both directly and through x_u.
Tolerances: FP32 Gram path 1e-4 relative to the oracle gradient; FP16 Gram
path 5e-3 relative (the 16-bit Gram perturbs A by ~1e-3 relative).
"""
import numpy as np
import pytest

import cuda_als_amd as ca
from cuda_als_amd import engine as E
from oracle import als_oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def subsample():
    """2,000 seeded users of the ml-latest-shaped workload, the item factors
    after 3 FP32 sweeps on the full matrix, and the users' held-out ratings."""
    d = ca.ml_latest_shape().generate()
    f = 64
    with E.Engine(E.make_config(f, 0.05, 3, E.SolveType.CHOL, E.Prec.FP32), d) as e:
        e.init_factors()
        e.train(3, with_rmse=False)
        V = e.get_factors(ca.ALS_SIDE_ITEMS)
    rng = np.random.default_rng(2000)
    users = np.sort(rng.choice(d.m, 2000, replace=False))
    rows = [np.arange(d.csr_ptr[u], d.csr_ptr[u + 1]) for u in users]
    ptr = np.concatenate([[0], np.cumsum([r.size for r in rows])]).astype(np.int32)
    cat = np.concatenate(rows)
    remap = -np.ones(d.m, np.int64)
    remap[users] = np.arange(users.size)
    tmask = remap[d.t_row] >= 0
    sub = ca.from_arrays(users.size, d.n, ptr, d.csr_idx[cat], d.csr_val[cat], remap[d.t_row[tmask]],
                         d.t_col[tmask], d.t_val[tmask])
    return sub, V.astype(np.float64), f


def _oracle_loss_and_grads(sub, V, lam, f):
    cfg = O.config(f, 0.0, O.SOLVE_CHOL)
    X, _ = O.update(cfg, sub.csr_ptr, sub.csr_idx, sub.csr_val, V, np.zeros((sub.m, f)), lam)
    pred = np.einsum("ij,ij->i", X[sub.t_row], V[sub.t_col])
    err = sub.t_val - pred
    N = err.size
    L = np.sqrt(np.mean(err ** 2))
    G = np.zeros_like(X)
    np.add.at(G, sub.t_row, (-err / (N * L))[:, None] * V[sub.t_col])
    Xb, lb, Yb = O.adjoint(f, sub.csr_ptr, sub.csr_idx, sub.csr_val, V, lam, X, G)
    np.add.at(Yb, sub.t_col, (-err / (N * L))[:, None] * X[sub.t_row])   # direct term
    return L, lb, Yb


def _gpu_loss_and_grads(sub, V, lam, f, prec):
    from cuda_als_amd.autograd import half_sweep
    dev = torch.device("cuda")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        eng = E.Engine(E.make_config(f, 0.05, 1, E.SolveType.CHOL, prec), sub)
        lam_t = torch.tensor(lam, dtype=torch.float32, device=dev, requires_grad=True)
        Y_t = torch.tensor(V, dtype=torch.float32, device=dev, requires_grad=True)
        X = half_sweep(lam_t, Y_t, eng)
        tr, tc = torch.tensor(sub.t_row, device=dev).long(), torch.tensor(sub.t_col, device=dev).long()
        tv = torch.tensor(sub.t_val, device=dev)
        pred = (X[tr] * Y_t[tc]).sum(1)
        L = torch.sqrt(torch.mean((tv - pred) ** 2))
        L.backward()
        out = (float(L.detach()), lam_t.grad.double().cpu().numpy(), Y_t.grad.double().cpu().numpy())
    torch.cuda.synchronize()
    eng.close()
    return out


@pytest.mark.parametrize("prec,tol", [(E.Prec.FP32, 1e-4), (E.Prec.FP16, 5e-3)])
def test_adjoint_matches_fp64_reference_gradient(subsample, prec, tol):
    sub, V, f = subsample
    lam = np.full(sub.m, 0.05)
    L0, lb0, Yb0 = _oracle_loss_and_grads(sub, V, lam, f)
    L1, lb1, Yb1 = _gpu_loss_and_grads(sub, V, lam, f, prec)
    print("loss", L0, L1, "lambda-grad rel err", np.abs(lb1 - lb0).max() / np.abs(lb0).max(),
          "Y-grad rel err", np.abs(Yb1 - Yb0).max() / np.abs(Yb0).max())
    assert abs(L1 - L0) <= tol * L0
    assert np.abs(lb1 - lb0).max() <= tol * np.abs(lb0).max()
    assert np.abs(Yb1 - Yb0).max() <= tol * np.abs(Yb0).max()


def test_adjoint_matches_central_differences(subsample):
    """FP64 central differences of the oracle loss (h = 1e-6) at a few
    coordinates of lambda and Y, against the GPU (FP32 Gram) gradient."""
    sub, V, f = subsample
    lam = np.full(sub.m, 0.05)
    _, lb, Yb = _gpu_loss_and_grads(sub, V, lam, f, E.Prec.FP32)
    cfg = O.config(f, 0.0, O.SOLVE_CHOL)

    def loss(lam_, V_):
        X, _ = O.update(cfg, sub.csr_ptr, sub.csr_idx, sub.csr_val, V_, np.zeros((sub.m, f)), lam_)
        pred = np.einsum("ij,ij->i", X[sub.t_row], V_[sub.t_col])
        return np.sqrt(np.mean((sub.t_val - pred) ** 2))

    h = 1e-6
    rng = np.random.default_rng(3)
    users = rng.choice(np.unique(sub.t_row), 5, replace=False)
    for u in users:
        e = np.zeros(sub.m); e[u] = h
        fd = (loss(lam + e, V) - loss(lam - e, V)) / (2 * h)
        assert abs(fd - lb[u]) <= 1e-3 * max(abs(fd), 1e-4 * np.abs(lb).max()), (u, fd, lb[u])
    items = rng.choice(np.unique(sub.t_col), 3, replace=False)
    for i in items:
        for j in rng.choice(f, 2, replace=False):
            E_ = np.zeros_like(V); E_[i, j] = h
            fd = (loss(lam, V + E_) - loss(lam, V - E_)) / (2 * h)
            assert abs(fd - Yb[i, j]) <= 1e-3 * max(abs(fd), 1e-4 * np.abs(Yb).max()), (i, j, fd, Yb[i, j])


def test_twenty_lambda_steps_reduce_heldout_rmse(subsample):
    """20 Adam steps (lr 0.1) on log(lambda_u) through the mixed-precision
    (FP16 Gram) forward and the adjoint lower the held-out RMSE."""
    from cuda_als_amd.autograd import half_sweep
    sub, V, f = subsample
    dev = torch.device("cuda")
    s = torch.cuda.Stream()
    hist = []
    with torch.cuda.stream(s):
        eng = E.Engine(E.make_config(f, 0.05, 1, E.SolveType.CHOL, E.Prec.FP16), sub)
        theta = torch.full((sub.m,), float(np.log(0.05)), device=dev, requires_grad=True)
        opt = torch.optim.Adam([theta], lr=0.1)
        Y = torch.tensor(V, dtype=torch.float32, device=dev)
        tr, tc = torch.tensor(sub.t_row, device=dev).long(), torch.tensor(sub.t_col, device=dev).long()
        tv = torch.tensor(sub.t_val, device=dev)
        for _ in range(20):
            opt.zero_grad()
            X = half_sweep(torch.exp(theta), Y, eng)
            L = torch.sqrt(torch.mean((tv - (X[tr] * Y[tc]).sum(1)) ** 2))
            L.backward()
            assert torch.isfinite(theta.grad).all()
            hist.append(float(L.detach()))
            opt.step()
    torch.cuda.synchronize()
    eng.close()
    print("held-out RMSE over 20 lambda steps:", [round(h, 5) for h in hist])
    assert hist[-1] < hist[0] - 1e-3
    assert hist[1] < hist[0]


def _random_ratings(m, n, lo, hi, n_test, seed):
    """m users with lo..hi ratings each over n items, integer stars; a test
    set of n_test random (user, item) pairs."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(lo, hi + 1, m)
    cols = np.concatenate([np.sort(rng.choice(n, c, replace=False)) for c in counts]).astype(np.int32)
    ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    vals = rng.integers(1, 6, cols.size).astype(np.float32)
    t_row = np.sort(rng.integers(0, m, n_test)).astype(np.int32)
    t_col = rng.integers(0, n, n_test).astype(np.int32)
    t_val = rng.integers(1, 6, n_test).astype(np.float32)
    return ca.from_arrays(m, n, ptr, cols, vals, t_row, t_col, t_val)


@pytest.mark.parametrize("f", [100, 143])
def test_adjoint_wide_factors(f):
    """f = 100 (T = 7) and f = 143 (T = 9, the widest system: the lane loops
    of the fixed-factor gradient cover factors 128..142 too) against the FP64
    reference gradient and central differences (h = 1e-6)."""
    sub = _random_ratings(240, 300, 20, 160, 3000, seed=f)
    rng = np.random.default_rng(f + 1)
    V = 0.3 * rng.standard_normal((sub.n, f))
    V = V.astype(np.float32).astype(np.float64)
    lam = np.full(sub.m, 0.05)
    L0, lb0, Yb0 = _oracle_loss_and_grads(sub, V, lam, f)
    L1, lb1, Yb1 = _gpu_loss_and_grads(sub, V, lam, f, E.Prec.FP32)
    print(f, "loss", L0, L1, "lambda-grad rel err", np.abs(lb1 - lb0).max() / np.abs(lb0).max(),
          "Y-grad rel err", np.abs(Yb1 - Yb0).max() / np.abs(Yb0).max())
    assert abs(L1 - L0) <= 1e-4 * L0
    assert np.abs(lb1 - lb0).max() <= 1e-4 * np.abs(lb0).max()
    assert np.abs(Yb1 - Yb0).max() <= 1e-4 * np.abs(Yb0).max()
    if f > 128:
        assert np.abs(Yb1[:, 128:]).max() > 0.1 * np.abs(Yb0[:, 128:]).max()
    cfg = O.config(f, 0.0, O.SOLVE_CHOL)

    def loss(lam_, V_):
        X, _ = O.update(cfg, sub.csr_ptr, sub.csr_idx, sub.csr_val, V_, np.zeros((sub.m, f)), lam_)
        pred = np.einsum("ij,ij->i", X[sub.t_row], V_[sub.t_col])
        return np.sqrt(np.mean((sub.t_val - pred) ** 2))

    h = 1e-6
    for u in rng.choice(np.unique(sub.t_row), 3, replace=False):
        e = np.zeros(sub.m); e[u] = h
        fd = (loss(lam + e, V) - loss(lam - e, V)) / (2 * h)
        assert abs(fd - lb1[u]) <= 1e-3 * max(abs(fd), 1e-4 * np.abs(lb1).max()), (u, fd, lb1[u])
    items = rng.choice(np.unique(sub.t_col), 2, replace=False)
    for i in items:
        for j in (0, f - 1, int(rng.integers(0, f))):
            E_ = np.zeros_like(V); E_[i, j] = h
            fd = (loss(lam, V + E_) - loss(lam, V - E_)) / (2 * h)
            assert abs(fd - Yb1[i, j]) <= 1e-3 * max(abs(fd), 1e-4 * np.abs(Yb1).max()), (i, j, fd, Yb1[i, j])


def test_chained_half_sweeps_keep_their_own_factorisation():
    """users then items in one autograd graph (U = solve(V), V2 = solve(U),
    loss on U . V2): each backward uses its own call's factorisation. The
    gradients with respect to lambda_u, lambda_i and V match the FP64 chain
    rule through two oracle adjoints. The engine-owned form refuses an
    adjoint of the side its saved factorisation does not belong to."""
    from cuda_als_amd.autograd import half_sweep
    f = 32
    d = _random_ratings(200, 150, 10, 60, 2000, seed=9)
    rng = np.random.default_rng(10)
    V = (0.3 * rng.standard_normal((d.n, f))).astype(np.float32).astype(np.float64)
    lu, li = np.full(d.m, 0.05), np.full(d.n, 0.08)
    # FP64 chain rule
    cfg = O.config(f, 0.0, O.SOLVE_CHOL)
    U, _ = O.update(cfg, d.csr_ptr, d.csr_idx, d.csr_val, V, np.zeros((d.m, f)), lu)
    V2, _ = O.update(cfg, d.csc_ptr, d.csc_idx, d.csc_val, U, np.zeros((d.n, f)), li)
    err = d.t_val - np.einsum("ij,ij->i", U[d.t_row], V2[d.t_col])
    N, L = err.size, np.sqrt(np.mean(err ** 2))
    w = (-err / (N * L))[:, None]
    gU = np.zeros_like(U); np.add.at(gU, d.t_row, w * V2[d.t_col])
    gV2 = np.zeros_like(V2); np.add.at(gV2, d.t_col, w * U[d.t_row])
    _, lbi, Ub = O.adjoint(f, d.csc_ptr, d.csc_idx, d.csc_val, U, li, V2, gV2)
    _, lbu, Vb = O.adjoint(f, d.csr_ptr, d.csr_idx, d.csr_val, V, lu, U, gU + Ub)
    dev = torch.device("cuda")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        eng = E.Engine(E.make_config(f, 0.05, 1, E.SolveType.CHOL, E.Prec.FP32), d)
        lu_t = torch.tensor(lu, dtype=torch.float32, device=dev, requires_grad=True)
        li_t = torch.tensor(li, dtype=torch.float32, device=dev, requires_grad=True)
        V_t = torch.tensor(V, dtype=torch.float32, device=dev, requires_grad=True)
        U_t = half_sweep(lu_t, V_t, eng, ca.ALS_SIDE_USERS)
        V2_t = half_sweep(li_t, U_t, eng, ca.ALS_SIDE_ITEMS)
        tr, tc = torch.tensor(d.t_row, device=dev).long(), torch.tensor(d.t_col, device=dev).long()
        pred = (U_t[tr] * V2_t[tc]).sum(1)
        Lg = torch.sqrt(torch.mean((torch.tensor(d.t_val, device=dev) - pred) ** 2))
        Lg.backward()
        g_lu, g_li, g_V = (t.grad.double().cpu().numpy() for t in (lu_t, li_t, V_t))
        # engine-owned factorisation: saved for the users side, adjoint of items refused
        Yp = torch.zeros(d.n, eng.info().ld, device=dev)
        Xp = torch.zeros(d.m, eng.info().ld, device=dev)
        eng.forward_dev(ca.ALS_SIDE_USERS, Yp.data_ptr(), Xp.data_ptr(), None, save=True)
        Y2 = torch.zeros(d.m, eng.info().ld, device=dev)
        X2 = torch.zeros(d.n, eng.info().ld, device=dev)
        lb = torch.zeros(d.n, device=dev)
        with pytest.raises(ca.AlsError):
            eng.adjoint_dev(ca.ALS_SIDE_ITEMS, Y2.data_ptr(), X2.data_ptr(), X2.data_ptr(), X2.data_ptr(),
                            lb.data_ptr(), None)
    torch.cuda.synchronize()
    eng.close()
    print("loss", L, float(Lg), "rel errs", np.abs(g_lu - lbu).max() / np.abs(lbu).max(),
          np.abs(g_li - lbi).max() / np.abs(lbi).max(), np.abs(g_V - Vb).max() / np.abs(Vb).max())
    assert abs(float(Lg) - L) <= 1e-4 * L
    assert np.abs(g_lu - lbu).max() <= 1e-4 * np.abs(lbu).max()
    assert np.abs(g_li - lbi).max() <= 1e-4 * np.abs(lbi).max()
    assert np.abs(g_V - Vb).max() <= 1e-4 * np.abs(Vb).max()
