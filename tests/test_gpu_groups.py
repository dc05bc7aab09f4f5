"""Several models over one rating matrix at the ml10M-shaped workload
(69,878 x 10,677, 10,000,054 ratings of which 1,000,006 are held out, row
sizes lognormal with sigma 1.3 and minimum 20, half-star ratings): f = 32,
64 and 128 with a BF16 Gram (FP32 accumulation), FP32 Cholesky and one
warm-started FP64-residual refinement step, updated as one group (one graph
launch per half-sweep, each model's kernels on a stream of their own).

Checked: every model's factors after three group sweeps equal its own
separate run bit for bit; one more half-sweep of each side from that state
matches, on sampled rows, the FP64 solve of the BF16-formed system within
5e-5 relative to max(1, |x|) (the warm-started refinement converges to
that system); the test RMSE is within 2e-4 of the FP64 oracle's exact solve
on the same state.
"""
import numpy as np
import pytest

import cuda_als_amd as ca
from cuda_als_amd import engine as E
from oracle import als_oracle as O

pytestmark = pytest.mark.gpu

FS = (32, 64, 128)


@pytest.fixture(scope="module")
def ml10m():
    return ca.ml10m_shape().generate()


def _engines(d, prec=E.Prec.BF16):
    return [E.Engine(E.make_config(f, 0.05, 1, E.SolveType.CHOL, prec, refine=1, refine_warm=True), d) for f in FS]


def test_ml10m_group_equals_separate_runs_and_matches_oracle(ml10m):
    d = ml10m
    sep, grp = _engines(d), _engines(d)
    for e in sep + grp:
        e.init_factors()
    with E.Group(grp) as g:
        for _ in range(3):
            g.sweep()
    for e in sep:
        for _ in range(3):
            e.sweep()
    rng = np.random.default_rng(10)
    cu, ci = np.diff(d.csr_ptr), np.diff(d.csc_ptr)
    users = np.unique(np.concatenate([rng.choice(d.m, 200, replace=False), np.argsort(cu)[-20:],
                                      rng.choice(np.where(cu < 48)[0], 60, replace=False)]))
    items = np.unique(np.concatenate([np.argsort(ci)[-20:], rng.choice(d.n, 60, replace=False)]))
    for es, eg, f in zip(sep, grp, FS):
        for side in (0, 1):
            a, b = es.get_factors(side), eg.get_factors(side)
            assert np.array_equal(a, b), (f, side, np.abs(a - b).max())
        info = eg.info()
        assert info.failures == 0
        V = eg.get_factors(ca.ALS_SIDE_ITEMS).astype(np.float64)
        eg.update(ca.ALS_SIDE_USERS)
        U1 = eg.get_factors(ca.ALS_SIDE_USERS)
        eg.update(ca.ALS_SIDE_ITEMS)
        V1 = eg.get_factors(ca.ALS_SIDE_ITEMS)
        rmse = eg.rmse()
        cfg = O.config(f, 0.05, O.SOLVE_CHOL, gram_prec=O.PREC_BF16)
        Uo = O.update_rows(cfg, users, d.csr_ptr, d.csr_idx, d.csr_val, V)
        Vo = O.update_rows(cfg, items, d.csc_ptr, d.csc_idx, d.csc_val, U1.astype(np.float64))
        eu = np.abs(U1[users] - Uo).max(axis=1) / np.maximum(1.0, np.abs(Uo).max(axis=1))
        ev = np.abs(V1[items] - Vo).max(axis=1) / np.maximum(1.0, np.abs(Vo).max(axis=1))
        # the exact system's solve of the same half-sweeps, for the test RMSE
        ex = O.config(f, 0.05, O.SOLVE_CHOL)
        Ux, _ = O.update(ex, d.csr_ptr, d.csr_idx, d.csr_val, V, np.zeros((d.m, f)))
        Vx, _ = O.update(ex, d.csc_ptr, d.csc_idx, d.csc_val, Ux, np.zeros((d.n, f)))
        rx = O.rmse(d.t_row, d.t_col, d.t_val, Ux, Vx)
        print(f"f={f}: dual rows {info.dual_rows[0]} users / {info.dual_rows[1]} items, split rows "
              f"{info.nsplit_rows[1]} items; sampled rows max rel err users {eu.max():.2e} items {ev.max():.2e}; "
              f"test RMSE {rmse:.6f} vs exact {rx:.6f}")
        assert eu.max() <= 5e-5 and ev.max() <= 5e-5
        assert abs(rmse - rx) <= 2e-4
    for e in sep + grp:
        e.close()


@pytest.mark.parametrize("f", [100, 128, 143])
def test_one_wave_kernel_variant_equals_default(f):
    """The update kernel's one-wave-per-SIMD variant (warm-started
    refinement, 7 to 9 tiles) runs the same code with the whole register
    file: bit-identical factors to the default two-wave kernel, on a
    netflix-shaped matrix (scaled) with light, whole and split rows.
    f = 128 also takes the software-pipelined Gram (16 gather segments),
    f = 143 the plain loop."""
    d = ca.scaled(ca.netflix_shape(), 0.03).generate()
    out = []
    for variant in (0, 1):
        cfg = E.make_config(f, 0.05, 2, E.SolveType.CHOL, E.Prec.FP16, refine=1, refine_warm=True, split_len=2048)
        with E.Engine(cfg, d) as e:
            info = e.info()
            assert info.grid_variant1 > 0
            for side in (0, 1):
                e.set_kernel_variant(side, variant)
            e.init_factors()
            e.train(2, with_rmse=False)
            out.append((e.get_factors(0), e.get_factors(1), e.info()))
    (U0, V0, i0), (U1, V1, i1) = out
    assert list(i1.kernel_variant) == [1, 1] and list(i0.kernel_variant) == [0, 0]
    assert i0.failures == 0 and i0.nsplit_rows[1] > 0
    assert np.array_equal(U0, U1) and np.array_equal(V0, V1)
    # and against the FP64 solve of the FP16-formed system, one users half-sweep from that state
    rng = np.random.default_rng(f)
    users = rng.choice(d.m, 200, replace=False)
    cfg = E.make_config(f, 0.05, 1, E.SolveType.CHOL, E.Prec.FP16, refine=1, refine_warm=True)
    with E.Engine(cfg, d) as e:
        e.set_kernel_variant(0, 1)
        e.set_factors(0, U1)
        e.set_factors(1, V1)
        e.update(ca.ALS_SIDE_USERS)
        Ug = e.get_factors(0)
    Uo = O.update_rows(O.config(f, 0.05, O.SOLVE_CHOL, gram_prec=O.PREC_FP16), users, d.csr_ptr, d.csr_idx,
                       d.csr_val, V1.astype(np.float64))
    err = np.abs(Ug[users] - Uo).max(axis=1) / np.maximum(1.0, np.abs(Uo).max(axis=1))
    print(f, "max rel err", err.max())
    assert err.max() <= 5e-5


@pytest.mark.parametrize("f,prec", [(100, E.Prec.FP16), (97, E.Prec.BF16), (120, E.Prec.FP16), (128, E.Prec.FP16)])
def test_one_wave_register_gathers_equal_default(f, prec):
    """The one-wave variant with the register-gather Gram (the Gram the
    autotuner picks for the items half): bit-identical factors to the
    default kernel with register gathers, after two warm-started sweeps of a
    netflix-shaped matrix (scaled) whose rows end on every chunk position
    (rows shorter than a chunk, one and two chunks, partial last chunks),
    for 7, 8 and 9 tiles and both 16-bit formats."""
    d = ca.scaled(ca.netflix_shape(), 0.03).generate()
    out = []
    for variant in (0, 1):
        cfg = E.make_config(f, 0.05, 1, E.SolveType.CHOL, prec, refine=1, refine_warm=True, split_len=2048)
        with E.Engine(cfg, d) as e:
            for side in (0, 1):
                e.set_kernel_variant(side, variant)
                e.set_gram_gather(side, 0)
            e.init_factors()
            e.train(2, with_rmse=False)
            out.append((e.get_factors(0), e.get_factors(1), e.info()))
    (U0, V0, i0), (U1, V1, i1) = out
    assert list(i1.kernel_variant) == [1, 1] and list(i1.gram_gather) == [0, 0]
    assert i0.failures == 0 and i1.failures == 0
    assert np.all(np.isfinite(U1)) and np.abs(U1).max() > 0
    assert np.array_equal(U0, U1) and np.array_equal(V0, V1)


@pytest.mark.parametrize("f,prec", [(100, E.Prec.FP16), (100, E.Prec.BF16), (64, E.Prec.FP16), (40, E.Prec.FP16)])
def test_gram_gather_modes_equal(f, prec):
    """The 16-bit Gram gathers either straight into LDS (LDS-DMA loads, two
    chunks in flight) or through registers: bit-identical factors after two
    warm-started sweeps of a netflix-shaped matrix (scaled) with split rows,
    light rows and partial last chunks, for 7, 5 and 3 tiles. Tile counts
    without the LDS-DMA Gram (f = 128) and the FP32 Gram refuse mode 1."""
    d = ca.scaled(ca.netflix_shape(), 0.03).generate()
    out = []
    for mode in (1, 0):
        cfg = E.make_config(f, 0.05, 2, E.SolveType.CHOL, prec, refine=1, refine_warm=True, split_len=1500)
        with E.Engine(cfg, d) as e:
            for side in (0, 1):
                e.set_gram_gather(side, mode)
            assert list(e.info().gram_gather) == [mode, mode]
            e.init_factors()
            e.train(2, with_rmse=False)
            info = e.info()
            assert info.failures == 0 and info.nsplit_rows[1] > 0
            out.append((e.get_factors(0), e.get_factors(1)))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    for ff, pp in ((128, E.Prec.FP16), (100, E.Prec.FP32)):
        with E.Engine(E.make_config(ff, 0.05, 1, E.SolveType.CHOL, pp), d) as e:
            with pytest.raises(ca.AlsError):
                e.set_gram_gather(0, 1)
            e.set_gram_gather(0, 0)
