"""A C++ caller written against the reference's own names (tests/abi_caller.cpp:
host_sparse_matrix's h_ arrays, the als_model constructor and enums,
train(), rsme_test(), h_UT / h_VT, save_libmf(), LU_solve_U with a batch
size and offset, updateXWithCGHost) compiles and links against the engine
library on the CPU, and on the GPU it reproduces the FP64 oracle on the
reference's toy_data workload (config/toy_data.ini: 15 x 13, f = 100,
lambda 0.05, 10 sweeps, SIMPLE Gram, LU solver): test RMSE and factors
within 1e-4, the LU_solve_U re-solve of every user within 1e-4 of the exact
solve, and CG within its exit threshold.
"""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import cuda_als_amd as ca
from cuda_als_amd import _capi as A

from .conftest import reference_toy

ROOT = Path(__file__).resolve().parents[1]
SRC = Path(__file__).resolve().parent / "abi_caller.cpp"


def build_caller(out: Path) -> subprocess.CompletedProcess:
    lib = A.LIB_PATH.parent
    cmd = ["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", f"-I{ROOT / 'include'}", "-I/opt/rocm/include",
           str(SRC), "-o", str(out), f"-L{lib}", "-lals_mi355x", f"-Wl,-rpath,{lib}", "-L/opt/rocm/lib",
           "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    return subprocess.run(cmd, capture_output=True, text=True, timeout=300)


def test_reference_named_caller_compiles_and_links(tmp_path):
    r = build_caller(tmp_path / "abi_caller")
    assert r.returncode == 0, r.stderr[-3000:]
    assert (tmp_path / "abi_caller").exists()


@pytest.mark.gpu
def test_reference_named_caller_matches_oracle(tmp_path):
    from oracle import als_oracle as O
    exe = tmp_path / "abi_caller"
    r = build_caller(exe)
    assert r.returncode == 0, r.stderr[-3000:]
    d, ini = reference_toy("toy_data")
    data = tmp_path / "data"
    data.mkdir()
    ca.save_bin(d, str(data))
    f, lam, iters = int(ini["als.f"]), float(ini["als.lambda"]), int(ini["als.als_iters"])
    model, users = tmp_path / "model.bin", tmp_path / "users.bin"
    args = [str(exe), str(data), ini["data.m"], ini["data.n"], ini["data.nnz_train"], ini["data.nnz_test"],
            ini["als.f"], ini["als.lambda"], ini["als.als_iters"], ini["als.als_calculate_vtvs_type"],
            ini["als.als_solve_type"], str(model), str(users)]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0
    kv = dict(re.findall(r"(\w+)=(\S+)", r.stdout))
    assert kv["solver"] == "LU" and kv["gram"] == "SIMPLE"
    U0, V0 = O.reference_init(d.m, d.n, f)
    Uo, Vo, ro, bad = O.train(O.config(f, lam, O.SOLVE_CHOL, store_fp32=True), d, iters, U0, V0)
    assert bad == 0
    assert abs(float(kv["rmse_test"]) - ro[-1]) <= 1e-4
    rows = d.train_rows()
    assert abs(float(kv["rmse_train"]) - O.rmse(rows, d.csr_idx, d.csr_val, Uo, Vo)) <= 1e-4
    raw = model.read_bytes()
    assert np.frombuffer(raw[:16], np.int32).tolist() == [0, d.m, d.n, f]
    body = np.frombuffer(raw[20:], np.float32)
    Um, Vm = body[:d.m * f].reshape(d.m, f), body[d.m * f:].reshape(d.n, f)
    print("factor err", np.abs(Um - Uo).max(), np.abs(Vm - Vo).max())
    assert np.abs(Um - Uo).max() <= 1e-4 and np.abs(Vm - Vo).max() <= 1e-4
    assert float(kv["h_UT0"]) == Um[0, 0] and float(kv["h_VT0"]) == Vm[0, 0]
    # LU_solve_U in two batches = every user solved exactly from the trained V
    Ug = np.fromfile(users, np.float32).reshape(d.m, f)
    Ux, _ = O.update(O.config(f, lam, O.SOLVE_CHOL), d.csr_ptr, d.csr_idx, d.csr_val, Vm.astype(np.float64),
                     np.zeros((d.m, f)))
    assert np.abs(Ug - Ux).max() <= 1e-4
    # 20 x 20 systems with lambda_min >= 20 and the exit at r.r < 1e-4: |x - x*| <= 1e-2 / 20
    assert float(kv["cg_max_abs_err"]) <= 1e-3
