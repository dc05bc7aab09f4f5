"""Batched CG on caller device buffers (als_cg_update_x / _fp16, the
reference's updateXWithCGHost / updateXWithCGHost_tt_fp16,
cumf_als_cg/cg.h:30-32) against the oracle's FP64 restatement of the same
recurrence (cumf_als_cg/cg.cu:37-232: warm start, iteration cap, exit once
r.r < tol after an update).

Tolerance: 1e-4 relative to max(1, |x|) for the FP32 recurrence against the
FP64 one over the same iterations (systems with condition numbers up to
~50, where six FP32 iterations stay within ~1e-6 of the FP64 ones).
"""
import numpy as np
import pytest
import torch

from cuda_als_amd import engine as E
from cuda_als_amd._capi import AlsError
from oracle import als_oracle as O

pytestmark = pytest.mark.gpu


def spd_batch(rng, batch, f):
    Y = rng.standard_normal((batch, 2 * f + 4, f)) / np.sqrt(2 * f + 4)
    A = np.einsum("bki,bkj->bij", Y, Y) + 0.5 * np.eye(f)[None]
    b = rng.standard_normal((batch, f))
    x0 = 0.1 * rng.standard_normal((batch, f))
    return A.astype(np.float32), b.astype(np.float32), x0.astype(np.float32)


def run_gpu(A, b, x0, iters, tol, half=False):
    dA = torch.from_numpy(A).cuda()
    if half:
        dA = dA.half()
    dx = torch.from_numpy(x0).cuda()
    E.cg_update_x(dA, dx, torch.from_numpy(b).cuda(), iters, tol)
    return dx.cpu().numpy()


def run_oracle(A, b, x0, iters, tol):
    return np.stack([O.cg_solve(A[i].astype(np.float64), b[i].astype(np.float64), x0[i].astype(np.float64),
                                iters, tol) for i in range(A.shape[0])])


def rel(a, ref):
    return (np.abs(a - ref).max(axis=1) / np.maximum(1.0, np.abs(ref).max(axis=1))).max()


@pytest.mark.parametrize("f", [1, 7, 16, 33, 100, 112, 128, 143])
def test_cg_batch_matches_oracle_recurrence(f):
    rng = np.random.default_rng(f)
    A, b, x0 = spd_batch(rng, 96, f)
    for iters, tol in ((6, 0.0), (6, 1e-4), (3, 1e-4)):
        xg = run_gpu(A, b, x0, iters, tol)
        xo = run_oracle(A, b, x0, iters, tol)
        print(f, iters, tol, rel(xg, xo))
        assert rel(xg, xo) <= 1e-4


@pytest.mark.parametrize("f", [7, 16, 40])
def test_cg_batch_converges_to_direct_solve(f):
    rng = np.random.default_rng(100 + f)
    A, b, x0 = spd_batch(rng, 32, f)
    xg = run_gpu(A, b, x0, 4 * f, 0.0)
    xd = np.linalg.solve(A.astype(np.float64), b.astype(np.float64)[..., None])[..., 0]
    assert rel(xg, xd) <= 1e-4


def test_cg_batch_exit_rule_and_warm_start():
    rng = np.random.default_rng(3)
    A, b, x0 = spd_batch(rng, 16, 24)
    # a huge threshold ends the loop after the first update: one CG step
    xg = run_gpu(A, b, x0, 6, 1e30)
    r = b - np.einsum("bij,bj->bi", A.astype(np.float64), x0)
    ar = np.einsum("bij,bj->bi", A.astype(np.float64), r)
    one = x0 + (np.einsum("bi,bi->b", r, r) / np.einsum("bi,bi->b", r, ar))[:, None] * r
    assert rel(xg, one) <= 1e-5
    # started at the solution, it stays there
    xs = np.linalg.solve(A.astype(np.float64), b.astype(np.float64)[..., None])[..., 0].astype(np.float32)
    assert rel(run_gpu(A, b, xs, 6, 1e-4), xs) <= 1e-5
    # zero iterations leave x alone
    assert np.array_equal(run_gpu(A, b, x0, 0, 1e-4), x0)


def test_cg_batch_fp16_matrix():
    rng = np.random.default_rng(5)
    A, b, x0 = spd_batch(rng, 64, 100)
    Ah = A.astype(np.float16).astype(np.float32)
    xg = run_gpu(A, b, x0, 6, 0.0, half=True)
    assert rel(xg, run_oracle(Ah, b, x0, 6, 0.0)) <= 1e-4


@pytest.mark.parametrize("f", [40, 100, 128, 143])
def test_cg_batch_fp32_matrix_rounded_to_fp16(f):
    """als_cg_update_x_tt_fp16 (FP32 A, each entry rounded to FP16 in the
    kernel): the same solutions as the FP16-A entry on A.half() (both round
    to nearest even), bit for bit, and the oracle's recurrence on the
    rounded A within the FP32 bar; f = 128 has nine tiles (one wave per
    SIMD), f = 143 takes the entry-by-entry loads."""
    import ctypes as C
    from cuda_als_amd._capi import lib
    rng = np.random.default_rng(f + 3)
    A, b, x0 = spd_batch(rng, 64, f)
    dA, db = torch.from_numpy(A).cuda(), torch.from_numpy(b).cuda()
    dx = torch.from_numpy(x0).cuda()
    rc = lib.als_cg_update_x_tt_fp16(C.c_void_p(dA.data_ptr()), C.c_void_p(dx.data_ptr()), C.c_void_p(db.data_ptr()),
                                     64, f, 6, 0.0, C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    xt = dx.cpu().numpy()
    assert np.array_equal(xt, run_gpu(A, b, x0, 6, 0.0, half=True))
    Ah = A.astype(np.float16).astype(np.float32)
    assert rel(xt, run_oracle(Ah, b, x0, 6, 0.0)) <= 1e-4


def test_cg_batch_unaligned_matrix_takes_the_entry_loads():
    """f % 4 == 0 with A 16-byte aligned loads a lane's four entries of a
    tile at once; an A one float off that alignment (a view into a larger
    buffer) takes the entry-by-entry loads: the same solutions."""
    rng = np.random.default_rng(11)
    A, b, x0 = spd_batch(rng, 48, 64)
    A = 0.5 * (A + A.transpose(0, 2, 1))   # exactly symmetric: both read orders see the same entries
    big = torch.zeros(A.size + 1, device="cuda")
    big[1:] = torch.from_numpy(A.ravel()).cuda()
    dA = big[1:].view(A.shape)
    assert dA.data_ptr() % 16 != 0
    dx = torch.from_numpy(x0).cuda()
    E.cg_update_x(dA, dx, torch.from_numpy(b).cuda(), 6, 0.0)
    assert np.array_equal(dx.cpu().numpy(), run_gpu(A, b, x0, 6, 0.0))


def test_cg_batch_rejects_bad_arguments():
    A = torch.zeros((2, 144, 144), device="cuda")
    x = torch.zeros((2, 144), device="cuda")
    with pytest.raises(AlsError):
        E.cg_update_x(A, x, x.clone())
    with pytest.raises(AlsError):
        E.cg_update_x(A[:, :4, :4].contiguous(), x[:, :4].contiguous(), x[:, :4].contiguous(), cg_iters=-1)
    # an empty batch is a no-op
    E.cg_update_x(A[:0, :4, :4], x[:0, :4], x[:0, :4])
