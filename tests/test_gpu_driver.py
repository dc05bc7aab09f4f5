"""The reference-compatible driver (cuda_als_amd/als_main) on a reference-
layout dataset and INI file, and the libmf model dump.

Reference behaviour followed: INI keys of main.cpp:28-47, .bin file names of
main.cpp:68-75, CSV event log of logger.cpp:71, libmf layout of
als_model.cu:2294-2312 (int 0, m, n, f, float mean |r|, U, V).
"""
import csv
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import cuda_als_amd as ca
from cuda_als_amd import engine as E

pytestmark = pytest.mark.gpu
MAIN = Path(__file__).resolve().parents[1] / "cuda_als_amd" / "als_main"


def write_ini(path, data, log, model, d, f, iters, vtvs, solve):
    path.write_text(f"""[general]
log_folder = {log}
als_runs = 1
model_path = {model}

[data]
data_folder = {data}
m = {d.m}
n = {d.n}
nnz_train = {d.nnz}
nnz_test = {d.n_test}

[als]
f = {f}
lambda = 0.05
als_iters = {iters}
; 0 SIMPLE ... 7 SMEM_ROW_MAJOR_TENSOR_SYMMETRIC_MULT_FRAG
als_calculate_vtvs_type = {vtvs}
als_solve_type = {solve}
m_batches = 2
n_batches = 2
""")


@pytest.mark.parametrize("vtvs,solve,prec,stype", [(0, 0, E.Prec.FP32, E.SolveType.LU),
                                                    (7, 2, E.Prec.FP16, E.SolveType.CUMF_ALS_CG_FP32)])
def test_driver_runs_reference_config(tmp_path, toy, vtvs, solve, prec, stype):
    assert MAIN.exists(), "build the driver: python cuda_als_amd/build.py"
    data, log = tmp_path / "data", tmp_path / "log"
    data.mkdir()
    log.mkdir()
    ca.save_bin(toy, str(data))
    ini = tmp_path / "toy.ini"
    model = tmp_path / "model.bin"
    f, iters = 16, 3
    write_ini(ini, data, log, model, toy, f, iters, vtvs, solve)
    r = subprocess.run([str(MAIN), str(ini)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    m = re.search(r"final RMSE test: ([0-9.]+)", r.stdout)
    assert m
    drv_rmse = float(m.group(1))

    # the same configuration through the engine
    with E.Engine(E.make_config(f, 0.05, iters, stype, prec), toy) as e:
        e.init_factors()
        e.train(iters, with_rmse=False)
        eng_rmse = e.rmse()
        U, V = e.get_factors(0), e.get_factors(1)
    assert abs(drv_rmse - eng_rmse) <= 2e-6

    # CSV event log: the reference's columns and events
    logs = list(log.glob("*_debug=false.csv"))
    assert len(logs) == 1
    rows = list(csv.reader(logs[0].open()))
    assert rows[0] == ["log_name", "id", "run_iter", "als_iter", "prev_ts", "curr_ts", "elapsed", "event_elapsed",
                       "text"]
    texts = [row[-1] for row in rows[1:]]
    for ev in ("ALS_UPDATE_U", "ALS_UPDATE_V", "ALS_ITER"):
        assert sum(t == f"event {ev} finished" for t in texts) == iters
    assert sum(t == "event MF_MODEL_TRAINING finished" for t in texts) == 1

    # libmf model: int 0, m, n, f, float mean |r|, U (m x f), V (n x f)
    raw = model.read_bytes()
    hdr = np.frombuffer(raw[:16], np.int32)
    assert hdr.tolist() == [0, toy.m, toy.n, f]
    mean = np.frombuffer(raw[16:20], np.float32)[0]
    assert abs(mean - np.abs(toy.csr_val).mean()) <= 1e-5
    body = np.frombuffer(raw[20:], np.float32)
    assert body.size == (toy.m + toy.n) * f
    Um, Vm = body[:toy.m * f].reshape(toy.m, f), body[toy.m * f:].reshape(toy.n, f)
    assert np.array_equal(Um, U) and np.array_equal(Vm, V)


def test_driver_reports_missing_files(tmp_path, toy):
    ini = tmp_path / "bad.ini"
    write_ini(ini, tmp_path / "nowhere", tmp_path, tmp_path / "m.bin", toy, 16, 1, 0, 0)
    r = subprocess.run([str(MAIN), str(ini)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "error" in r.stderr
