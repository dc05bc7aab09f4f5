"""Host-side logic and the C ABI surface (CPU only, no device calls).

Golden data: the reference's toy datasets (tests/golden/*.json, transcribed
from its data-prep logs by tests/golden/make_toy_fixtures.py) pin the
CSR/CSC/COO layout and the CSR -> CSC transpose.
"""
import ctypes as C
import json
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import cuda_als_amd as ca
from cuda_als_amd import _capi as A

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = Path(__file__).resolve().parent / "golden"


def load_toy(name):
    d = json.loads((GOLDEN / f"{name}.json").read_text())
    a = {k: np.array(v) for k, v in d["arrays"].items()}
    return d["m"], d["n"], a


@pytest.mark.parametrize("name", ["toy_data", "toy_data_2"])
def test_transpose_reproduces_reference_csc(name):
    m, n, a = load_toy(name)
    cp, ci, cv = ca.csr_to_csc(m, n, a["R_train_csr.indptr"], a["R_train_csr.indices"], a["R_train_csr.data"])
    assert np.array_equal(cp, a["R_train_csc.indptr"])
    assert np.array_equal(ci, a["R_train_csc.indices"])
    assert np.array_equal(cv, a["R_train_csc.data"].astype(np.float32))


@pytest.mark.parametrize("name", ["toy_data", "toy_data_2"])
def test_bin_round_trip_in_reference_layout(tmp_path, name):
    m, n, a = load_toy(name)
    d = ca.from_arrays(m, n, a["R_train_csr.indptr"], a["R_train_csr.indices"], a["R_train_csr.data"],
                       a["R_test_coo.row"], a["R_test_coo.col"], a["R_test_coo.data"])
    ca.save_bin(d, str(tmp_path))
    # raw layout: int32 / float32 little endian, file names of main.cpp:68-75
    assert np.array_equal(np.fromfile(tmp_path / "R_train_csc.indptr.bin", np.int32), a["R_train_csc.indptr"])
    assert np.array_equal(np.fromfile(tmp_path / "R_test_coo.data.bin", np.float32), a["R_test_coo.data"])
    e = ca.load_bin(str(tmp_path), m, n, d.nnz, d.n_test)
    for k in ("csr_ptr", "csr_idx", "csr_val", "csc_ptr", "csc_idx", "csc_val", "t_row", "t_col", "t_val"):
        assert np.array_equal(getattr(d, k), getattr(e, k))


def test_bin_loader_rejects_missing_and_wrong_size(tmp_path):
    """cuda_sparse_matrix.cpp:9-21: missing or wrong-size files are errors."""
    m, n, a = load_toy("toy_data")
    with pytest.raises(ca.AlsError) as ei:
        ca.load_bin(str(tmp_path), m, n, 36, 10)
    assert ei.value.code == -3
    d = ca.from_arrays(m, n, a["R_train_csr.indptr"], a["R_train_csr.indices"], a["R_train_csr.data"],
                       a["R_test_coo.row"], a["R_test_coo.col"], a["R_test_coo.data"])
    ca.save_bin(d, str(tmp_path))
    with pytest.raises(ca.AlsError) as ei:
        ca.load_bin(str(tmp_path), m, n, 37, 10)   # wrong nnz
    assert "size" in str(ei.value)


def test_synth_toy_workload_shape():
    d = ca.toy_1k().generate()
    assert (d.m, d.n, d.nnz, d.n_test) == (1000, 500, 18000, 2000)
    assert set(np.unique(d.csr_val)) <= {1.0, 2.0, 3.0, 4.0, 5.0}
    # distinct (user, item) pairs over train + test
    keys = np.concatenate([d.train_rows().astype(np.int64) * d.n + d.csr_idx, d.t_row.astype(np.int64) * d.n + d.t_col])
    assert np.unique(keys).size == keys.size
    # items ascending in every CSR row; test COO sorted by (row, col)
    for u in range(d.m):
        s, e = d.csr_ptr[u], d.csr_ptr[u + 1]
        assert np.all(np.diff(d.csr_idx[s:e]) > 0)
    tk = d.t_row.astype(np.int64) * d.n + d.t_col
    assert np.all(np.diff(tk) > 0)
    # item popularity decreases with the index (Zipf)
    cnt = np.diff(d.csc_ptr)
    assert cnt[:10].sum() > cnt[-10:].sum() * 5


def test_synth_is_deterministic_and_hits_exact_totals():
    cfg = ca.scaled(ca.ml10m_shape(), 0.05)
    a, b = cfg.generate(), cfg.generate()
    assert np.array_equal(a.csr_idx, b.csr_idx) and np.array_equal(a.csr_val, b.csr_val)
    assert a.nnz + a.n_test == cfg.total and a.n_test == cfg.n_test
    vals = np.unique(a.csr_val)
    assert np.all(np.isclose(vals * 2, np.round(vals * 2))) and vals.min() >= 0.5 and vals.max() <= 5.0
    c = ca.SynthConfig(**{**cfg.__dict__, "seed": 43}).generate()
    assert not np.array_equal(a.csr_idx, c.csr_idx)


def test_lognormal_rows_respect_bounds():
    cfg = ca.scaled(ca.netflix_shape(), 0.02)
    d = cfg.generate()
    full = np.diff(d.csr_ptr) + np.bincount(d.t_row, minlength=d.m)
    assert full.min() >= cfg.row_min and full.max() <= min(cfg.row_max, d.n)
    assert full.sum() == cfg.total


def test_schedule_splits_heavy_rows_and_covers_every_rating():
    rng = np.random.default_rng(0)
    sizes = np.concatenate([rng.integers(0, 50, 500), [5000, 12000]])
    ptr = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    st = A.ScheduleStats()
    units = np.zeros((2000, 4), np.int32)
    rc = A.lib.als_schedule_preview(sizes.size, ptr.ctypes.data_as(C.c_void_p), 1000, 0, 64, A.ALS_PREC_FP16,
                                    C.byref(st), units.ctypes.data_as(C.c_void_p), units.shape[0])
    assert rc == 0
    assert st.nsplit_rows == 2 and st.nslots == 5 + 12
    u = units[:st.nunits]
    covered = np.zeros(ptr[-1], np.int32)
    for row, kb, ke, sid in u:
        assert ptr[row] <= kb <= ke <= ptr[row + 1]
        covered[kb:ke] += 1
    assert np.all(covered == 1)
    # every row appears (empty rows too, so they get x = 0)
    assert set(u[:, 0]) == set(range(sizes.size))
    # heaviest first
    cost = (u[:, 2] - u[:, 1])
    assert cost[0] >= cost[-1]


def test_config_from_reference_mapping():
    cfg = A.AlsConfig()
    assert A.lib.als_config_from_reference(80, 0.048, 10, 7, 2, 32, 2, 2, C.byref(cfg)) == 0
    assert (cfg.f, cfg.iters, cfg.solve_type, cfg.gram_prec, cfg.m_batches) == (80, 10, 2, A.ALS_PREC_FP16, 2)
    assert abs(cfg.lambda_ - 0.048) < 1e-7
    assert A.lib.als_config_from_reference(100, 0.05, 10, 0, 0, 32, 1, 1, C.byref(cfg)) == 0
    assert cfg.gram_prec == A.ALS_PREC_FP32 and cfg.solve_type == A.ALS_SOLVE_LU
    assert A.lib.als_config_from_reference(100, 0.05, 10, 8, 0, 32, 1, 1, C.byref(cfg)) != 0
    assert A.lib.als_config_from_reference(100, 0.05, 10, 0, 3, 32, 1, 1, C.byref(cfg)) != 0


def _declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(als_[a-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    names = _declared_functions()
    assert {"als_create", "als_update", "als_adjoint_dev", "als_synth_generate", "als_schedule_preview"} <= names
    out = subprocess.run(["nm", "-D", "--defined-only", str(A.LIB_PATH)], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = names - exported
    assert not missing, missing
    for n in names:
        getattr(A.lib, n)


def test_library_exports_reference_cpp_entry_points():
    """The cuMF CG host functions with the reference's C++ signatures
    (cumf_als_cg/cg.h:30,32, declared in include/als_cg.h) and the C++ mirror
    of als_model / host_sparse_matrix (include/als_model.hpp)."""
    out = subprocess.run(["nm", "-DC", "--defined-only", str(A.LIB_PATH)], capture_output=True, text=True).stdout
    for sig in ("updateXWithCGHost(float*, float*, float*, int, int, float)",
                "updateXWithCGHost_tt_fp16(float*, float*, float*, int, int, float)",
                "als_model::LU_solve_U(int, int)", "als_model::LU_solve_V(int, int)", "als_model::train(",
                "als_model::rsme_test()", "als_model::rsme_train()", "als_model::save_libmf(char const*)",
                "host_sparse_matrix::load_csr(", "host_sparse_matrix::load_csc(", "host_sparse_matrix::load_coo(",
                "host_sparse_matrix::malloc_load_file("):
        assert sig in out, sig


def test_engine_creation_fails_cleanly_without_a_device():
    from cuda_als_amd import engine as E
    from tests.conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(ca.AlsError):
        E.Engine(E.make_config(16))
