"""Rating matrices in the reference layout and the seeded workload generator.

The arrays follow the reference's binary layout (cuda_sparse_matrix.h:8-38):
CSR of users x items (int32 row_ptr, int32 col_idx, float32 val), its CSC
(the CSR of R^T) and a COO test set sorted by (row, col).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from ._capi import ALS_SYNTH_LOGNORMAL_ROWS, ALS_SYNTH_UNIFORM_PAIRS, AlsError, SynthSpec, lib


@dataclass
class RatingData:
    m: int
    n: int
    csr_ptr: np.ndarray
    csr_idx: np.ndarray
    csr_val: np.ndarray
    csc_ptr: np.ndarray
    csc_idx: np.ndarray
    csc_val: np.ndarray
    t_row: np.ndarray
    t_col: np.ndarray
    t_val: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.csr_idx.size)

    @property
    def n_test(self) -> int:
        return int(self.t_row.size)

    def train_rows(self) -> np.ndarray:
        return np.repeat(np.arange(self.m, dtype=np.int32), np.diff(self.csr_ptr))


def _check(rc: int) -> None:
    if rc != 0:
        raise AlsError(rc, lib.als_data_last_error().decode())


def _from_handle(h: C.c_void_p) -> RatingData:
    m, n, nnz, nt = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
    _check(lib.als_dataset_info(h, C.byref(m), C.byref(n), C.byref(nnz), C.byref(nt)))
    m, n, nnz, nt = m.value, n.value, nnz.value, nt.value

    def arr(p, count, dtype):
        if count == 0:
            return np.zeros(0, dtype=dtype)
        return np.ctypeslib.as_array(p, shape=(count,)).astype(dtype, copy=True)

    ip, ii, fv = C.POINTER(C.c_int)(), C.POINTER(C.c_int)(), C.POINTER(C.c_float)()
    _check(lib.als_dataset_csr(h, C.byref(ip), C.byref(ii), C.byref(fv)))
    csr = (arr(ip, m + 1, np.int32), arr(ii, nnz, np.int32), arr(fv, nnz, np.float32))
    _check(lib.als_dataset_csc(h, C.byref(ip), C.byref(ii), C.byref(fv)))
    csc = (arr(ip, n + 1, np.int32), arr(ii, nnz, np.int32), arr(fv, nnz, np.float32))
    _check(lib.als_dataset_test(h, C.byref(ip), C.byref(ii), C.byref(fv)))
    te = (arr(ip, nt, np.int32), arr(ii, nt, np.int32), arr(fv, nt, np.float32))
    return RatingData(m, n, *csr, *csc, *te)


def _ptr(a: np.ndarray) -> C.c_void_p:
    return a.ctypes.data_as(C.c_void_p)


@dataclass
class SynthConfig:
    """Seeded synthetic workload (see cuda_als_amd/csrc/als_data.cpp)."""
    m: int
    n: int
    total: int
    n_test: int
    row_model: int = ALS_SYNTH_LOGNORMAL_ROWS
    ln_mu: float = 0.0
    ln_sigma: float = 1.0
    row_min: int = 1
    row_max: int = 1 << 30
    zipf_s: float = 1.0
    rank: int = 8
    noise_sigma: float = 0.5
    rating_step: float = 1.0
    rating_min: float = 1.0
    rating_max: float = 5.0
    rating_offset: float = 3.0
    seed: int = 42

    def generate(self) -> RatingData:
        spec = SynthSpec(self.m, self.n, self.total, self.n_test, self.row_model, self.ln_mu, self.ln_sigma,
                         self.row_min, min(self.row_max, self.n), self.zipf_s, self.rank, self.noise_sigma,
                         self.rating_step, self.rating_min, self.rating_max, self.rating_offset, self.seed)
        h = C.c_void_p()
        _check(lib.als_synth_generate(C.byref(spec), C.byref(h)))
        try:
            return _from_handle(h)
        finally:
            lib.als_dataset_free(h)


def from_arrays(m: int, n: int, csr_ptr, csr_idx, csr_val, t_row, t_col, t_val) -> RatingData:
    """Dataset from a CSR train matrix and COO test set (CSC derived)."""
    a = [np.ascontiguousarray(x, dtype=t) for x, t in ((csr_ptr, np.int32), (csr_idx, np.int32),
                                                       (csr_val, np.float32), (t_row, np.int32),
                                                       (t_col, np.int32), (t_val, np.float32))]
    h = C.c_void_p()
    _check(lib.als_dataset_from_csr(m, n, a[1].size, _ptr(a[0]), _ptr(a[1]), _ptr(a[2]), a[3].size,
                                    _ptr(a[3]), _ptr(a[4]), _ptr(a[5]), C.byref(h)))
    try:
        return _from_handle(h)
    finally:
        lib.als_dataset_free(h)


def load_bin(folder: str, m: int, n: int, nnz_train: int, nnz_test: int) -> RatingData:
    """Reference .bin loader (cuda_sparse_matrix.cpp:6-69, names main.cpp:68-75)."""
    h = C.c_void_p()
    _check(lib.als_dataset_load_bin(folder.encode(), m, n, nnz_train, nnz_test, C.byref(h)))
    try:
        return _from_handle(h)
    finally:
        lib.als_dataset_free(h)


def save_bin(d: RatingData, folder: str) -> None:
    h = C.c_void_p()
    _check(lib.als_dataset_from_csr(d.m, d.n, d.nnz, _ptr(d.csr_ptr), _ptr(d.csr_idx), _ptr(d.csr_val), d.n_test,
                                    _ptr(d.t_row), _ptr(d.t_col), _ptr(d.t_val), C.byref(h)))
    try:
        _check(lib.als_dataset_save_bin(h, folder.encode()))
    finally:
        lib.als_dataset_free(h)


def csr_to_csc(rows: int, cols: int, ptr, idx, val):
    ptr, idx, val = (np.ascontiguousarray(ptr, np.int32), np.ascontiguousarray(idx, np.int32),
                     np.ascontiguousarray(val, np.float32))
    cp = np.zeros(cols + 1, np.int32)
    ci = np.zeros(idx.size, np.int32)
    cv = np.zeros(idx.size, np.float32)
    _check(lib.als_csr_to_csc(rows, cols, idx.size, _ptr(ptr), _ptr(idx), _ptr(val), _ptr(cp), _ptr(ci), _ptr(cv)))
    return cp, ci, cv


# ----------------------------------------------------------------- workloads
# The five workloads the engine is specified and measured on. `total`
# counts train + test ratings; seeds are fixed at 42.

def toy_1k() -> SynthConfig:
    """1,000 users x 500 items, 20,000 ratings (2,000 held out); items Zipf(1.1),
    users uniform; integer stars from a rank-8 model + N(0, 0.5^2)."""
    return SynthConfig(m=1000, n=500, total=20000, n_test=2000, row_model=ALS_SYNTH_UNIFORM_PAIRS, zipf_s=1.1,
                       rank=8, noise_sigma=0.5)


def netflix_shape() -> SynthConfig:
    """480,189 x 17,770, 100,480,507 ratings of which 1,408,395 are held out
    (99,072,112 train); row nnz lognormal(ln 96, 1.2484) clipped to
    [1, 17000]; items Zipf(1.0) without replacement; integer stars from a
    rank-20 model + N(0, 0.8^2)."""
    return SynthConfig(m=480189, n=17770, total=100480507, n_test=1408395, ln_mu=math.log(96.0),
                       ln_sigma=1.2484, row_min=1, row_max=17000, zipf_s=1.0, rank=20, noise_sigma=0.8)


def ml10m_shape() -> SynthConfig:
    """69,878 x 10,677, 10,000,054 ratings (1,000,006 held out); row nnz
    lognormal(4.119, 1.3) with minimum 20; half stars from a rank-16 model
    + N(0, 0.9^2)."""
    return SynthConfig(m=69878, n=10677, total=10000054, n_test=1000006, ln_mu=4.119, ln_sigma=1.3, row_min=20,
                       row_max=10677, zipf_s=1.0, rank=16, noise_sigma=0.9, rating_step=0.5, rating_min=0.5,
                       rating_max=5.0)


def ml_latest_shape() -> SynthConfig:
    """283,228 x 58,098, 27,753,444 ratings (2,775,344 held out); row nnz
    lognormal(ln 30, 1.5386); half stars from a rank-16 model + N(0, 0.9^2)."""
    return SynthConfig(m=283228, n=58098, total=27753444, n_test=2775344, ln_mu=math.log(30.0), ln_sigma=1.5386,
                       row_min=1, row_max=58098, zipf_s=1.0, rank=16, noise_sigma=0.9, rating_step=0.5,
                       rating_min=0.5, rating_max=5.0)


def scaled(cfg: SynthConfig, factor: float, seed: int | None = None) -> SynthConfig:
    """Same distributions at a fraction of the size (parity tests)."""
    m = max(8, int(cfg.m * factor))
    n = max(8, int(cfg.n * factor))
    total = max(16, int(cfg.total * factor * factor))
    # keep the row-size law, cap rows at the item count
    out = SynthConfig(**{**cfg.__dict__})
    out.m, out.n, out.total = m, n, min(total, m * n // 2)
    out.n_test = max(1, int(cfg.n_test / cfg.total * out.total))
    out.row_max = min(cfg.row_max, n)
    out.row_min = min(cfg.row_min, max(1, out.total // m))
    if cfg.row_model == ALS_SYNTH_LOGNORMAL_ROWS:
        mean_row = out.total / m
        out.ln_mu = math.log(max(1.0, mean_row)) - cfg.ln_sigma ** 2 / 2
    if seed is not None:
        out.seed = seed
    return out
