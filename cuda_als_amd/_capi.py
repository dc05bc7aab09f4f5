"""ctypes declarations of the engine's C ABI (include/als.h, include/als_data.h).

The library must have been built (python cuda_als_amd/build.py); there is no
fallback implementation: a missing library raises at import of this module.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

# One HIP runtime per process. PyTorch-ROCm ships its own libamdhip64 (SONAME
# libamdhip64.so.7) and links it as "libamdhip64.so". If torch is loaded first,
# the engine's NEEDED libamdhip64.so.7 binds to that same copy; the other
# order loads a second runtime, which then sees no device. So torch, when it
# is installed, is loaded before the engine.
try:
    import torch  # noqa: F401
except ImportError:
    pass

LIB_PATH = Path(__file__).resolve().parent / "libals_mi355x.so"

ALS_OK = 0
ALS_SOLVE_LU, ALS_SOLVE_CG_FP16, ALS_SOLVE_CG_FP32, ALS_SOLVE_CHOL = 0, 1, 2, 3
ALS_PREC_FP32, ALS_PREC_BF16, ALS_PREC_FP16 = 0, 1, 2
ALS_SIDE_USERS, ALS_SIDE_ITEMS = 0, 1
ALS_RMSE_TEST, ALS_RMSE_TRAIN = 0, 1
ALS_INIT_REFERENCE, ALS_INIT_ZERO = 0, 1
ALS_SYNTH_UNIFORM_PAIRS, ALS_SYNTH_LOGNORMAL_ROWS = 0, 1

STATUS_NAMES = {0: "ALS_OK", -1: "ALS_E_INVALID", -2: "ALS_E_HIP", -3: "ALS_E_IO", -4: "ALS_E_STATE",
                -5: "ALS_E_NUMERIC", -6: "ALS_E_UNSUPPORTED"}


class AlsConfig(C.Structure):
    _fields_ = [("f", C.c_int), ("lambda_", C.c_float), ("iters", C.c_int), ("solve_type", C.c_int),
                ("gram_prec", C.c_int), ("cg_iters", C.c_int), ("cg_tol", C.c_float),
                ("refine_steps", C.c_int), ("m_batches", C.c_int), ("n_batches", C.c_int),
                ("device", C.c_int), ("batch_work", C.c_int), ("split_len", C.c_int), ("use_graph", C.c_int),
                ("refine_warm", C.c_int)]


class AlsInfo(C.Structure):
    _fields_ = [("T", C.c_int), ("ld", C.c_int), ("grid", C.c_int * 2), ("lds_bytes", C.c_int),
                ("nbatches", C.c_int * 2), ("nunits", C.c_int * 2), ("nsplit_rows", C.c_int * 2),
                ("split_len", C.c_int64 * 2), ("failures", C.c_int64), ("m", C.c_int64), ("n", C.c_int64),
                ("nnz", C.c_int64), ("n_test", C.c_int64), ("dual_max", C.c_int), ("dual_rows", C.c_int * 2),
                ("kernel_variant", C.c_int * 2), ("grid_variant1", C.c_int),
                ("dual_limit", C.c_int * 2), ("gram_gather", C.c_int * 2)]


class ScheduleStats(C.Structure):
    _fields_ = [("nunits", C.c_int64), ("nsplit_rows", C.c_int64), ("nbatches", C.c_int64), ("nslots", C.c_int64),
                ("split_len", C.c_int64), ("batch_work", C.c_int64)]


class TuneResult(C.Structure):
    _fields_ = [("split_len", C.c_int64 * 2), ("batch_work", C.c_int64 * 2), ("default_ms", C.c_float * 2),
                ("best_ms", C.c_float * 2), ("candidates", C.c_int * 2), ("restage_max", C.c_int * 2),
                ("kernel_variant", C.c_int * 2), ("grid", C.c_int * 2),
                ("dual_limit", C.c_int * 2), ("gram_gather", C.c_int * 2)]


class SynthSpec(C.Structure):
    _fields_ = [("m", C.c_int64), ("n", C.c_int64), ("total", C.c_int64), ("n_test", C.c_int64),
                ("row_model", C.c_int), ("ln_mu", C.c_double), ("ln_sigma", C.c_double),
                ("row_min", C.c_int64), ("row_max", C.c_int64), ("zipf_s", C.c_double), ("rank", C.c_int),
                ("noise_sigma", C.c_double), ("rating_step", C.c_double), ("rating_min", C.c_double),
                ("rating_max", C.c_double), ("rating_offset", C.c_double), ("seed", C.c_uint64)]


P = C.c_void_p
I32P = C.POINTER(C.c_int)
F32P = C.POINTER(C.c_float)
F64P = C.POINTER(C.c_double)

# name: (restype, argtypes)
SIGNATURES = {
    # als.h
    "als_config_init": (C.c_int, [C.POINTER(AlsConfig)]),
    "als_config_from_reference": (C.c_int, [C.c_int, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.c_int, C.POINTER(AlsConfig)]),
    "als_create": (C.c_int, [C.POINTER(AlsConfig), C.POINTER(P)]),
    "als_destroy": (None, [P]),
    "als_last_error": (C.c_char_p, [P]),
    "als_set_stream": (C.c_int, [P, P]),
    "als_set_train": (C.c_int, [P, C.c_int64, C.c_int64, C.c_int64, P, P, P, P, P, P]),
    "als_set_test": (C.c_int, [P, C.c_int64, P, P, P]),
    "als_init_factors": (C.c_int, [P, C.c_int]),
    "als_set_factors": (C.c_int, [P, C.c_int, P, C.c_int64, C.c_int]),
    "als_get_factors": (C.c_int, [P, C.c_int, P, C.c_int64, C.c_int]),
    "als_update": (C.c_int, [P, C.c_int]),
    "als_update_rows": (C.c_int, [P, C.c_int, C.c_int64, C.c_int64, C.c_int]),
    "als_sweep": (C.c_int, [P]),
    "als_train": (C.c_int, [P, C.c_int, P]),
    "als_rmse": (C.c_int, [P, C.c_int, F64P]),
    "als_synchronize": (C.c_int, [P]),
    "als_get_info": (C.c_int, [P, C.POINTER(AlsInfo)]),
    "als_save_libmf": (C.c_int, [P, C.c_char_p]),
    "als_device_factors": (C.c_int, [P, C.c_int, C.POINTER(P)]),
    "als_forward_dev": (C.c_int, [P, C.c_int, P, P, P, C.c_int]),
    "als_adjoint_dev": (C.c_int, [P, C.c_int, P, P, P, P, P, P]),
    "als_factor_save_bytes": (C.c_int, [P, C.c_int, C.POINTER(C.c_int64)]),
    "als_forward_dev_ws": (C.c_int, [P, C.c_int, P, P, P, P, C.c_int64]),
    "als_adjoint_dev_ws": (C.c_int, [P, C.c_int, P, P, P, C.c_int64, P, P, P, P]),
    "als_schedule_preview": (C.c_int, [C.c_int64, P, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(ScheduleStats), P, C.c_int64]),
    "als_autotune": (C.c_int, [P, C.c_int, C.POINTER(TuneResult)]),
    "als_set_schedule": (C.c_int, [P, C.c_int, C.c_int64, C.c_int64, C.c_int]),
    "als_set_kernel_variant": (C.c_int, [P, C.c_int, C.c_int]),
    "als_set_grid": (C.c_int, [P, C.c_int, C.c_int]),
    "als_set_dual_limit": (C.c_int, [P, C.c_int, C.c_int]),
    "als_set_gram_gather": (C.c_int, [P, C.c_int, C.c_int]),
    "als_group_create": (C.c_int, [C.POINTER(P), C.c_int, C.POINTER(P)]),
    "als_group_destroy": (None, [P]),
    "als_group_last_error": (C.c_char_p, [P]),
    "als_group_update": (C.c_int, [P, C.c_int]),
    "als_group_sweep": (C.c_int, [P]),
    "als_cg_update_x": (C.c_int, [P, P, P, C.c_int, C.c_int, C.c_int, C.c_float, P]),
    "als_cg_update_x_fp16": (C.c_int, [P, P, P, C.c_int, C.c_int, C.c_int, C.c_float, P]),
    "als_cg_update_x_tt_fp16": (C.c_int, [P, P, P, C.c_int, C.c_int, C.c_int, C.c_float, P]),
    "als_rmse_dev_noalloc": (C.c_int, [P, P, P, C.c_int64, P, P, C.c_int, C.c_int, P, P]),
    "als_noalloc_last_error": (C.c_char_p, []),
    # als_data.h
    "als_synth_generate": (C.c_int, [C.POINTER(SynthSpec), C.POINTER(P)]),
    "als_dataset_from_csr": (C.c_int, [C.c_int64, C.c_int64, C.c_int64, P, P, P, C.c_int64, P, P, P, C.POINTER(P)]),
    "als_dataset_load_bin": (C.c_int, [C.c_char_p, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.POINTER(P)]),
    "als_dataset_save_bin": (C.c_int, [P, C.c_char_p]),
    "als_dataset_info": (C.c_int, [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int64)]),
    "als_dataset_csr": (C.c_int, [P, C.POINTER(I32P), C.POINTER(I32P), C.POINTER(F32P)]),
    "als_dataset_csc": (C.c_int, [P, C.POINTER(I32P), C.POINTER(I32P), C.POINTER(F32P)]),
    "als_dataset_test": (C.c_int, [P, C.POINTER(I32P), C.POINTER(I32P), C.POINTER(F32P)]),
    "als_dataset_train_rows": (C.c_int, [P, C.POINTER(I32P)]),
    "als_dataset_free": (None, [P]),
    "als_csr_to_csc": (C.c_int, [C.c_int64, C.c_int64, C.c_int64, P, P, P, P, P, P]),
    "als_data_last_error": (C.c_char_p, []),
}


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: build it with `python cuda_als_amd/build.py`")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class AlsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code
