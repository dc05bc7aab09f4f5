"""MI355X-native ALS engine: the per-row normal-equation update of
vserdyuk/cuda_als (als_model update-U / update-V, RMSE) as hand-written
gfx950 HIP kernels behind a C ABI (include/als.h).

Importing this package loads libals_mi355x.so; there is no CPU fallback.
"""
from ._capi import (ALS_INIT_REFERENCE, ALS_INIT_ZERO, ALS_RMSE_TEST, ALS_RMSE_TRAIN, ALS_SIDE_ITEMS,
                    ALS_SIDE_USERS, AlsError)
from .data import (RatingData, SynthConfig, csr_to_csc, from_arrays, load_bin, ml10m_shape, ml_latest_shape,
                   netflix_shape, save_bin, scaled, toy_1k)
from .engine import AlsModel, CalculateVtvsType, Engine, Prec, SolveType, make_config

__all__ = [
    "ALS_INIT_REFERENCE", "ALS_INIT_ZERO", "ALS_RMSE_TEST", "ALS_RMSE_TRAIN", "ALS_SIDE_ITEMS", "ALS_SIDE_USERS",
    "AlsError", "AlsModel", "CalculateVtvsType", "Engine", "Prec", "RatingData", "SolveType", "SynthConfig",
    "csr_to_csc", "from_arrays", "load_bin", "make_config", "ml10m_shape", "ml_latest_shape", "netflix_shape",
    "save_bin", "scaled", "toy_1k",
]
