"""Python view of the engine's C ABI (include/als.h).

`Engine` wraps one als_engine handle. `AlsModel` mirrors the reference's
als_model (als_model.h:29-43): same constructor arguments, train(),
rsme_test(), rsme_train(), save_libmf() and h_UT / h_VT after training.
"""
from __future__ import annotations

import ctypes as C
from enum import IntEnum

import numpy as np

from . import _capi as A
from ._capi import AlsConfig, AlsError, AlsInfo, lib
from .data import RatingData


class CalculateVtvsType(IntEnum):   # als_model.h:12-21
    SIMPLE = 0
    SMEM_ROW_MAJOR = 1
    SMEM_COL_MAJOR = 2
    SMEM_ROW_MAJOR_NO_CALC = 3
    SMEM_COL_MAJOR_TWO_THREADS = 4
    SMEM_ROW_MAJOR_TENSOR = 5
    SMEM_ROW_MAJOR_TENSOR_SYMMETRIC = 6
    SMEM_ROW_MAJOR_TENSOR_SYMMETRIC_MULT_FRAG = 7


class SolveType(IntEnum):           # als_model.h:23-27, plus CHOL
    LU = 0
    CUMF_ALS_CG_FP16 = 1
    CUMF_ALS_CG_FP32 = 2
    CHOL = 3


class Prec(IntEnum):
    FP32 = 0
    BF16 = 1
    FP16 = 2


def _p(a: np.ndarray) -> C.c_void_p:
    return a.ctypes.data_as(C.c_void_p)


def make_config(f: int, lambda_: float = 0.05, iters: int = 10, solve: int = SolveType.CHOL,
                prec: int = Prec.FP32, cg_iters: int = 6, cg_tol: float = 1e-4, refine: int = 0,
                use_graph: bool = False, batch_work: int = 0, split_len: int = 0, device: int = 0,
                refine_warm: bool = False) -> AlsConfig:
    cfg = AlsConfig()
    lib.als_config_init(C.byref(cfg))
    cfg.f, cfg.lambda_, cfg.iters, cfg.solve_type, cfg.gram_prec = f, lambda_, iters, int(solve), int(prec)
    cfg.cg_iters, cfg.cg_tol, cfg.refine_steps = cg_iters, cg_tol, refine
    cfg.use_graph, cfg.batch_work, cfg.split_len, cfg.device = int(use_graph), batch_work, split_len, device
    cfg.refine_warm = int(refine_warm)
    return cfg


class Engine:
    def __init__(self, cfg: AlsConfig, data: RatingData | None = None):
        self.cfg = cfg
        h = C.c_void_p()
        rc = lib.als_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise AlsError(rc, "als_create failed (no HIP device, or unsupported configuration)")
        self.h = h
        self.m = self.n = 0
        if data is not None:
            self.set_data(data)

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise AlsError(rc, lib.als_last_error(self.h).decode())

    def close(self) -> None:
        # at interpreter exit the module's globals may already be gone
        if getattr(self, "h", None) and lib is not None:
            lib.als_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- data / factors
    def set_data(self, d: RatingData) -> None:
        self._keep = d
        self._check(lib.als_set_train(self.h, d.m, d.n, d.nnz, _p(d.csr_ptr), _p(d.csr_idx), _p(d.csr_val),
                                      _p(d.csc_ptr), _p(d.csc_idx), _p(d.csc_val)))
        self._check(lib.als_set_test(self.h, d.n_test, _p(d.t_row), _p(d.t_col), _p(d.t_val)))
        self.m, self.n = d.m, d.n

    def set_stream(self, stream_handle: int | None) -> None:
        self._check(lib.als_set_stream(self.h, C.c_void_p(stream_handle or 0)))

    def init_factors(self, mode: int = A.ALS_INIT_REFERENCE) -> None:
        self._check(lib.als_init_factors(self.h, mode))

    def set_factors(self, side: int, x: np.ndarray) -> None:
        x = np.ascontiguousarray(x, dtype=np.float32)
        self._check(lib.als_set_factors(self.h, side, _p(x), x.shape[0], x.shape[1]))

    def get_factors(self, side: int) -> np.ndarray:
        rows = self.m if side == A.ALS_SIDE_USERS else self.n
        out = np.zeros((rows, self.cfg.f), np.float32)
        self._check(lib.als_get_factors(self.h, side, _p(out), rows, self.cfg.f))
        return out

    # -- compute
    def update(self, side: int) -> None:
        self._check(lib.als_update(self.h, side))

    def sweep(self) -> None:
        self._check(lib.als_sweep(self.h))

    def train(self, sweeps: int = -1, with_rmse: bool = True) -> np.ndarray | None:
        n = self.cfg.iters if sweeps < 0 else sweeps
        out = np.zeros(max(n, 1), np.float64) if with_rmse else None
        self._check(lib.als_train(self.h, n, _p(out) if out is not None else None))
        return out[:n] if out is not None else None

    def rmse(self, which: int = A.ALS_RMSE_TEST) -> float:
        v = C.c_double()
        self._check(lib.als_rmse(self.h, which, C.byref(v)))
        return v.value

    def synchronize(self) -> None:
        self._check(lib.als_synchronize(self.h))

    def info(self) -> AlsInfo:
        i = AlsInfo()
        self._check(lib.als_get_info(self.h, C.byref(i)))
        return i

    def autotune(self, sides: int = 3) -> dict:
        """Times candidate schedules for the selected sides (bit 0 users, bit 1
        items) and keeps the fastest; factors are left unchanged."""
        r = A.TuneResult()
        self._check(lib.als_autotune(self.h, sides, C.byref(r)))
        return {"split_len": list(r.split_len), "batch_work": list(r.batch_work),
                "default_ms": list(r.default_ms), "best_ms": list(r.best_ms), "candidates": list(r.candidates),
                "restage_max": list(r.restage_max), "kernel_variant": list(r.kernel_variant),
                "grid": list(r.grid), "dual_limit": list(r.dual_limit), "gram_gather": list(r.gram_gather)}

    def set_schedule(self, side: int, split_len: int = 0, batch_work: int = 0, restage_max: int = -1) -> None:
        self._check(lib.als_set_schedule(self.h, side, split_len, batch_work, restage_max))

    def set_kernel_variant(self, side: int, variant: int) -> None:
        self._check(lib.als_set_kernel_variant(self.h, side, variant))

    def set_grid(self, side: int, grid: int) -> None:
        self._check(lib.als_set_grid(self.h, side, grid))

    def set_dual_limit(self, side: int, limit: int) -> None:
        self._check(lib.als_set_dual_limit(self.h, side, limit))

    def set_gram_gather(self, side: int, mode: int) -> None:
        self._check(lib.als_set_gram_gather(self.h, side, mode))

    def device_factors(self, side: int) -> int:
        p = C.c_void_p()
        self._check(lib.als_device_factors(self.h, side, C.byref(p)))
        return p.value or 0

    def save_libmf(self, path: str) -> None:
        self._check(lib.als_save_libmf(self.h, path.encode()))

    # -- differentiable half-sweep (device pointers)
    def forward_dev(self, side: int, d_Y: int, d_X: int, d_lambda: int | None, save: bool = True) -> None:
        self._check(lib.als_forward_dev(self.h, side, C.c_void_p(d_Y), C.c_void_p(d_X),
                                        C.c_void_p(d_lambda or 0), int(save)))

    def adjoint_dev(self, side: int, d_Y: int, d_X: int, d_gbar: int, d_xbar: int, d_lbar: int,
                    d_ybar: int | None) -> None:
        self._check(lib.als_adjoint_dev(self.h, side, C.c_void_p(d_Y), C.c_void_p(d_X), C.c_void_p(d_gbar),
                                        C.c_void_p(d_xbar), C.c_void_p(d_lbar), C.c_void_p(d_ybar or 0)))

    def factor_save_bytes(self, side: int) -> int:
        v = C.c_int64()
        self._check(lib.als_factor_save_bytes(self.h, side, C.byref(v)))
        return v.value

    def forward_dev_ws(self, side: int, d_Y: int, d_X: int, d_lambda: int | None, d_save: int,
                       save_bytes: int) -> None:
        self._check(lib.als_forward_dev_ws(self.h, side, C.c_void_p(d_Y), C.c_void_p(d_X), C.c_void_p(d_lambda or 0),
                                           C.c_void_p(d_save), save_bytes))

    def adjoint_dev_ws(self, side: int, d_Y: int, d_X: int, d_save: int, save_bytes: int, d_gbar: int, d_xbar: int,
                       d_lbar: int, d_ybar: int | None) -> None:
        self._check(lib.als_adjoint_dev_ws(self.h, side, C.c_void_p(d_Y), C.c_void_p(d_X), C.c_void_p(d_save),
                                           save_bytes, C.c_void_p(d_gbar), C.c_void_p(d_xbar), C.c_void_p(d_lbar),
                                           C.c_void_p(d_ybar or 0)))


def cg_update_x(A, x, b, cg_iters: int = 6, tol: float = 1e-4) -> None:
    """Batched CG on device tensors, the reference's updateXWithCGHost
    (cumf_als_cg/cg.h:30; FP16 A: updateXWithCGHost_tt_fp16, cg.h:32).
    A: (batch, f, f) float32 or float16, symmetric positive definite; x:
    (batch, f) float32, warm start in and solution out; b: (batch, f)
    float32; all contiguous on the current HIP device. Runs on torch's
    current stream and returns when it has finished, like the reference."""
    import torch
    if A.dim() != 3 or A.shape[1] != A.shape[2] or x.shape != A.shape[:2] or b.shape != A.shape[:2]:
        raise ValueError("A must be (batch, f, f), x and b (batch, f)")
    if x.dtype != torch.float32 or b.dtype != torch.float32 or A.dtype not in (torch.float32, torch.float16):
        raise ValueError("x, b float32; A float32 or float16")
    if not (A.is_cuda and x.is_cuda and b.is_cuda) or not (A.is_contiguous() and x.is_contiguous()
                                                          and b.is_contiguous()):
        raise ValueError("A, x, b must be contiguous device tensors")
    fn = lib.als_cg_update_x if A.dtype == torch.float32 else lib.als_cg_update_x_fp16
    rc = fn(C.c_void_p(A.data_ptr()), C.c_void_p(x.data_ptr()), C.c_void_p(b.data_ptr()), int(A.shape[0]),
            int(A.shape[1]), int(cg_iters), float(tol), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    if rc != 0:
        raise AlsError(rc, "als_cg_update_x failed (f in 1..143, cg_iters >= 0, tol >= 0)")


class Group:
    """Several engines (models of different f) over one rating matrix,
    updated by one graph launch per half-sweep, each model's kernels on a
    stream of their own (als_group_*). The engines stay usable on their own
    and must outlive the group."""

    def __init__(self, engines: list[Engine]):
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e.h for e in self.engines])
        h = C.c_void_p()
        rc = lib.als_group_create(arr, len(self.engines), C.byref(h))
        if rc != 0:
            raise AlsError(rc, "als_group_create failed (1 to 4 distinct engines over the same rating matrix "
                               "and device)")
        self.h = h

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise AlsError(rc, lib.als_group_last_error(self.h).decode())

    def update(self, side: int) -> None:
        self._check(lib.als_group_update(self.h, side))

    def sweep(self) -> None:
        self._check(lib.als_group_sweep(self.h))

    def close(self) -> None:
        if getattr(self, "h", None) and lib is not None:
            lib.als_group_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class AlsModel:
    """Mirror of the reference's als_model (als_model.h:29-43)."""

    CALCULATE_VTVS_TYPE = CalculateVtvsType
    SOLVE_TYPE = SolveType

    def __init__(self, train: RatingData, test: RatingData | None, f: int, lambda_: float, iters: int,
                 calculate_vtvs_type: int = CalculateVtvsType.SIMPLE, solve_type: int = SolveType.LU,
                 smem_col_cnt: int = 32, m_batches: int = 1, n_batches: int = 1):
        cfg = AlsConfig()
        rc = lib.als_config_from_reference(f, lambda_, iters, int(calculate_vtvs_type), int(solve_type),
                                           smem_col_cnt, m_batches, n_batches, C.byref(cfg))
        if rc != 0:
            raise AlsError(rc, "invalid als_model configuration")
        if test is not None and test is not train:
            train = RatingData(train.m, train.n, train.csr_ptr, train.csr_idx, train.csr_val, train.csc_ptr,
                               train.csc_idx, train.csc_val, test.t_row, test.t_col, test.t_val)
        self.engine = Engine(cfg, train)
        self.m, self.n, self.f, self.lambda_, self.iters = train.m, train.n, f, lambda_, iters
        self.h_UT = np.zeros((self.m, f), np.float32)
        self.h_VT = np.zeros((self.n, f), np.float32)

    def train(self) -> None:
        self.engine.init_factors(A.ALS_INIT_REFERENCE)
        self.engine.train(self.iters, with_rmse=False)
        self.h_UT = self.engine.get_factors(A.ALS_SIDE_USERS)
        self.h_VT = self.engine.get_factors(A.ALS_SIDE_ITEMS)

    def rsme_test(self) -> float:
        return self.engine.rmse(A.ALS_RMSE_TEST)

    def rsme_train(self) -> float:
        return self.engine.rmse(A.ALS_RMSE_TRAIN)

    def save_libmf(self, path: str) -> None:
        self.engine.save_libmf(path)
