"""als_rmse_dev_noalloc: the RMSE kernels on caller device buffers with
their partial sums in the library's fixed device array (no allocation).

Checked against als_rmse (the engine's entry point, same kernels, partial
sums in engine memory) on the same factors and ratings: the sum of squared
errors is bit-identical, on two calls in a row with the second smaller,
and on replays of a captured HIP graph holding the call. Collected last:
it must pass alone and after every other test.
"""
import ctypes as C
import math

import numpy as np
import pytest

import cuda_als_amd as ca
from cuda_als_amd import _capi as A
from cuda_als_amd import engine as E

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

fn = A.lib.als_rmse_dev_noalloc


def _call(t_row, t_col, t_val, n, dU, dV, f, ld, out, stream):
    return fn(C.c_void_p(t_row.data_ptr()), C.c_void_p(t_col.data_ptr()), C.c_void_p(t_val.data_ptr()), n,
              C.c_void_p(dU), C.c_void_p(dV), f, ld, C.c_void_p(out.data_ptr()), C.c_void_p(stream))


def test_noalloc_rmse_equals_engine_rmse(toy):
    f = 100
    rng = np.random.default_rng(31)
    U = (0.2 * rng.standard_normal((toy.m, f))).astype(np.float32)
    V = (0.2 * rng.standard_normal((toy.n, f))).astype(np.float32)
    dev = torch.device("cuda")
    with E.Engine(E.make_config(f), toy) as e:
        e.set_factors(0, U)
        e.set_factors(1, V)
        ld = e.info().ld
        dU, dV = e.device_factors(0), e.device_factors(1)
        rows = torch.tensor(toy.t_row, device=dev)
        cols = torch.tensor(toy.t_col, device=dev)
        vals = torch.tensor(toy.t_val, device=dev)
        out = torch.zeros(2, dtype=torch.float64, device=dev)
        s = torch.cuda.Stream()
        results = []
        for n in (toy.n_test, toy.n_test // 2 + 3):       # the second call smaller
            tr, tc, tv = (np.ascontiguousarray(a[:n]) for a in (toy.t_row, toy.t_col, toy.t_val))
            assert A.lib.als_set_test(e.h, n, tr.ctypes.data, tc.ctypes.data, tv.ctypes.data) == 0
            want = e.rmse(ca.ALS_RMSE_TEST)
            torch.cuda.synchronize()
            with torch.cuda.stream(s):
                out.fill_(float("nan"))
                assert _call(rows, cols, vals, n, dU, dV, f, ld, out, s.cuda_stream) == 0, \
                    A.lib.als_noalloc_last_error()
            torch.cuda.synchronize()
            total, rmse = out.cpu().numpy()
            print(n, want, math.sqrt(total / n), rmse)
            assert math.sqrt(total / n) == want
            assert abs(rmse - want) <= 4e-16 * want
            results.append(total)
        assert results[0] != results[1]
        # sizes and layouts it cannot take are refused before any launch
        assert _call(rows, cols, vals, -1, dU, dV, f, ld, out, s.cuda_stream) == -1   # ALS_E_INVALID
        assert _call(rows, cols, vals, 10, dU, dV, f, ld - 2, out, s.cuda_stream) == -1
        assert _call(rows, cols, vals, 10, dU, dV, f, f - 1, out, s.cuda_stream) == -1
        assert b"ld" in A.lib.als_noalloc_last_error()


def test_noalloc_rmse_in_captured_graph(toy):
    """The call captured in a HIP graph and replayed: every replay rewrites
    the output with the eager value (the output is set to NaN and the
    factors of the engine are touched by eager work between replays)."""
    f = 48
    rng = np.random.default_rng(32)
    U = (0.2 * rng.standard_normal((toy.m, f))).astype(np.float32)
    V = (0.2 * rng.standard_normal((toy.n, f))).astype(np.float32)
    dev = torch.device("cuda")
    with E.Engine(E.make_config(f), toy) as e:
        e.set_factors(0, U)
        e.set_factors(1, V)
        want = e.rmse(ca.ALS_RMSE_TEST)
        ld, dU, dV = e.info().ld, e.device_factors(0), e.device_factors(1)
        rows = torch.tensor(toy.t_row, device=dev)
        cols = torch.tensor(toy.t_col, device=dev)
        vals = torch.tensor(toy.t_val, device=dev)
        out = torch.zeros(2, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            rc = _call(rows, cols, vals, toy.n_test, dU, dV, f, ld, out, torch.cuda.current_stream().cuda_stream)
        assert rc == 0, A.lib.als_noalloc_last_error()
        for rep in range(3):
            out.fill_(float("nan"))
            e.set_factors(1, V)          # eager engine work (same values) between replays
            g.replay()
            torch.cuda.synchronize()
            total = float(out[0].cpu())
            assert math.sqrt(total / toy.n_test) == want, rep
