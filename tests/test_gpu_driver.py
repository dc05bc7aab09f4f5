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


@pytest.mark.parametrize("name", ["toy_data", "toy_data_2"])
def test_reference_toy_configs_match_oracle(tmp_path, name):
    """The reference's own runnable configurations (config/toy_data.ini:
    15 x 13, f = 100; config/toy_data_2.ini: 32 x 23, f = 20; both SIMPLE
    Gram, LU solver, lambda 0.05, 10 sweeps, 10 runs) on their prepared
    arrays (the prep-log golden fixtures), through the driver: the test RMSE
    of every sweep (the CALC_ITER_RMSE lines and the final one) of every run
    within 1e-4 of the FP64 oracle. Factors after 10 sweeps: with one
    FP64-residual refinement step ([als] refine_steps = 1) within 1e-4; as
    configured (plain FP32 solve) within 1.5x the distance the same ALS done
    in float32 by numpy/LAPACK has from the oracle (1.3e-4 and 1.4e-4 on
    these two: on these heavily underdetermined rows the sweeps amplify
    FP32 rounding), and 1e-4 where that is larger."""
    from oracle import als_oracle as O

    from .conftest import reference_toy
    d, ini = reference_toy(name)
    data, log = tmp_path / "data", tmp_path / "log"
    data.mkdir()
    log.mkdir()
    ca.save_bin(d, str(data))
    f, lam, iters, runs = (int(ini["als.f"]), float(ini["als.lambda"]), int(ini["als.als_iters"]),
                           int(ini["general.als_runs"]))
    U0, V0 = O.reference_init(d.m, d.n, f)
    Uo, Vo, ro, bad = O.train(O.config(f, lam, O.SOLVE_CHOL, store_fp32=True), d, iters, U0, V0)
    assert bad == 0

    def run(conf, label):
        sections = {}
        for k, v in conf.items():
            sec, key = k.split(".", 1)
            sections.setdefault(sec, []).append(f"{key} = {v}")
        text = "".join(f"[{s}]\n" + "\n".join(lines) + "\n\n" for s, lines in sections.items())
        p = tmp_path / f"{label}.ini"
        p.write_text(text)
        r = subprocess.run([str(MAIN), str(p)], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        return r.stdout

    conf = dict(ini, **{"general.log_folder": str(log), "data.data_folder": str(data)})
    out = run(conf, "ref")
    per_sweep = [float(x) for x in re.findall(r"RMSE test: ([0-9.]+)", out)]
    assert len(per_sweep) == runs * iters     # iters - 1 per-sweep lines + the final line, per run
    got = np.array(per_sweep).reshape(runs, iters)
    print(name, "driver", got[0], "\noracle", ro)
    assert np.abs(got - ro[None, :]).max() <= 1e-4
    # single runs with the model dump: the factors
    from .test_gpu_parity import lapack_fp32_train
    Ul, Vl, _ = lapack_fp32_train(d, f, iters, lam)
    lapack = max(np.abs(Ul - Uo).max(), np.abs(Vl - Vo).max())
    for refine, bound in (("0", max(1e-4, 1.5 * lapack)), ("1", 1e-4)):
        model = tmp_path / f"model{refine}.bin"
        run(dict(conf, **{"general.als_runs": "1", "general.model_path": str(model), "als.refine_steps": refine}),
            f"dump{refine}")
        body = np.frombuffer(model.read_bytes()[20:], np.float32)
        Um, Vm = body[:d.m * f].reshape(d.m, f), body[d.m * f:].reshape(d.n, f)
        err = max(np.abs(Um - Uo).max(), np.abs(Vm - Vo).max())
        print(f"refine_steps {refine}: factor err {err:.3e} (float32 LAPACK: {lapack:.3e})")
        assert err <= bound
