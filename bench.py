"""Headline benchmark: full ALS sweeps/s at the netflix-shaped workload.

Workload: 480,189 users x 17,770 items, 99,072,112 train ratings (1,408,395
held out), f = 100, lambda = 0.05, seeded synthetic data (seed 42). A step is
one full sweep (all users from V, then all items from U) with the data
already resident in HBM, timed as replays of the captured sweep graph. The
mixed-precision path (FP16 Gram on MFMA with FP32 accumulation, FP32
Cholesky, FP64-residual refinement) is the measured line; the plain FP32 path
(FP32 MFMA Gram, FP32 Cholesky) is timed beside it for the speed ratio, and
so is the reference's production solver, FP32 CG with 6 iterations. Each
path's test RMSE over 10 sweeps from the reference init is compared with the
FP64 oracle's trajectory of the same workload (tests/golden fixture).

    python bench.py [--gpus 1] [--steps 10] [--warmup 3]

Prints one JSON line. `roofline` covers the dominant kernel (the per-row
update kernel of the half-sweep that takes longest), timed with HIP events
on the engine's stream; `cpu_baseline` is the FP64 oracle on a row sample.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "ALS full sweeps/s @100M nnz f=100, mixed-prec vs FP32 at equal test RMSE"
PEAK_FP16_TFLOPS = 2500.0     # MI355X dense FP16/BF16 MFMA
PEAK_FP32_TFLOPS = 157.3      # MI355X FP32 matrix
PEAK_HBM_TBS = 8.0            # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse_traffic(kernel_tag: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary."""
    p = ROOT / "profiles" / "traffic_update_netflix_shape.md"
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        m = re.match(rf"^traffic_bytes_per_launch\s+{kernel_tag}\s+(\S+)", line)
        if m:
            return float(m.group(1))
    return None


def time_path(torch, E, ca, d, cfg, steps, warmup, tag, autotune=True, oracle=None):
    """Warmup, then `steps` sweeps timed as whole-sweep graph replays (the
    headline figure), then `steps` sweeps launched half by half with HIP
    events between the halves on the engine's stream (per-half times). With
    `oracle` (the FP64 trajectory fixture and the name of its solver), a
    last pass from the reference init compares the test RMSE of every sweep
    with it. Returns a dict."""
    stream = torch.cuda.Stream()          # default priority; a real (non-null) stream handle
    torch.cuda.set_stream(stream)
    t = time.perf_counter()
    eng = E.Engine(cfg, d)
    torch.cuda.synchronize()
    upload_s = time.perf_counter() - t
    eng.set_stream(stream.cuda_stream)
    eng.init_factors()
    tune = None
    if autotune:
        tune = eng.autotune()
        log(f"[{tag}] autotune: {tune}")
    for _ in range(warmup):
        eng.sweep()                       # the first one captures the sweep graph
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.sweep()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    for s in range(steps):
        ev[s][0].record(stream)
        eng.update(ca.ALS_SIDE_USERS)
        ev[s][1].record(stream)
        eng.update(ca.ALS_SIDE_ITEMS)
        ev[s][2].record(stream)
    torch.cuda.synchronize()
    users_ms = sum(e[0].elapsed_time(e[1]) for e in ev) / steps
    items_ms = sum(e[1].elapsed_time(e[2]) for e in ev) / steps
    rmse = eng.rmse(ca.ALS_RMSE_TEST)
    info = eng.info()
    log(f"[{tag}] {steps} sweeps (graph) in {dt:.3f}s -> {steps / dt:.2f} sweeps/s; half by half: users "
        f"{users_ms:.3f} ms, items {items_ms:.3f} ms; test RMSE {rmse:.6f}; failures {info.failures}")
    out = {"eng": eng, "dt": dt, "users_ms": users_ms, "items_ms": items_ms, "rmse": rmse, "info": info,
           "upload_s": upload_s, "tune": tune, "vs_oracle": None}
    if oracle is not None:
        fix, name = oracle
        eng.init_factors()
        got = []
        for _ in range(fix["sweeps"]):
            eng.sweep()
            got.append(eng.rmse(ca.ALS_RMSE_TEST))
        want = fix[f"rmse_{name}"]
        out["vs_oracle"] = {"oracle": f"FP64 {name}", "sweeps": fix["sweeps"],
                            "max_abs_delta": max(abs(a - b) for a, b in zip(got, want)),
                            "final_delta": got[-1] - want[-1]}
        log(f"[{tag}] test RMSE vs the FP64 oracle ({name}), per sweep: "
            + " ".join(f"{a - b:+.1e}" for a, b in zip(got, want)))
    return out


def group_leg(torch, E, ca, steps, warmup, autotune=True):
    """Three models (f = 32, 64, 128; BF16 Gram, FP32 Cholesky + one warm-started
    refinement step) over the ml10M-shaped workload, updated as one group (one
    graph launch per half-sweep), against the same three engines run one after
    the other on one stream and run concurrently on their own streams."""
    d = ca.ml10m_shape().generate()
    fs = (32, 64, 128)
    s0 = torch.cuda.Stream()
    torch.cuda.set_stream(s0)
    engs = [E.Engine(E.make_config(f, 0.05, steps, E.SolveType.CHOL, E.Prec.BF16, refine=1, refine_warm=True), d)
            for f in fs]
    for e in engs:
        e.init_factors()
        if autotune:
            e.autotune()
    own = [torch.cuda.Stream() for _ in engs]

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return steps / (time.perf_counter() - t0)

    out = {"workload": f"ml10M-shape {d.m}x{d.n}, {d.nnz} train ratings; models f={list(fs)}, bf16 Gram, "
                       "fp32 Cholesky + 1 warm-started fp64-residual refinement step", "unit": "sweeps/s (all models)"}
    for e in engs:
        e.set_stream(s0.cuda_stream)
    with E.Group(engs) as g:
        out["group"] = timed(g.sweep)
    out["separate_one_stream"] = timed(lambda: [e.sweep() for e in engs])
    for e, st in zip(engs, own):
        e.set_stream(st.cuda_stream)
    out["separate_own_streams"] = timed(lambda: [e.sweep() for e in engs])
    out["per_model_alone"] = {}
    for e, f, st in zip(engs, fs, own):
        out["per_model_alone"][f"f{f}"] = timed(e.sweep)
    for e in engs:
        e.close()
    log(f"[group] {out}")
    return out


def adjoint_leg(torch, E, ca, steps, warmup):
    """The differentiable half-sweep at the ml-latest shape (283,228 users x
    58,098 items, 24,978,100 train ratings, f = 64, per-user lambda, FP16
    Gram): the forward users half-sweep keeping every row's Cholesky factor
    (als_forward_dev_ws) and the hand-written adjoint on it
    (als_adjoint_dev_ws: x-bar, d/d lambda_u, d/dY), each timed with HIP
    events on the engine's stream. The adjoint's floor is reading the saved
    factors once from HBM."""
    import torch.nn.functional as F
    d = ca.ml_latest_shape().generate()
    f = 64
    s = torch.cuda.Stream()
    torch.cuda.set_stream(s)
    with E.Engine(E.make_config(f, 0.05, 1, E.SolveType.CHOL, E.Prec.FP16), d) as e:
        e.set_stream(s.cuda_stream)
        e.init_factors()
        ld = e.info().ld
        dev = torch.device("cuda")
        Y = F.pad(torch.tensor(e.get_factors(ca.ALS_SIDE_ITEMS), device=dev), (0, ld - f)).contiguous()
        X = torch.zeros(d.m, ld, device=dev)
        lam = torch.full((d.m,), 0.05, device=dev)
        nbytes = e.factor_save_bytes(ca.ALS_SIDE_USERS)
        Ls = torch.empty(nbytes // 4, device=dev)
        g = torch.randn(d.m, ld, device=dev) * 1e-3
        g[:, f:] = 0
        xbar, lbar, ybar = torch.zeros_like(X), torch.zeros_like(lam), torch.zeros_like(Y)

        def fwd():
            e.forward_dev_ws(ca.ALS_SIDE_USERS, Y.data_ptr(), X.data_ptr(), lam.data_ptr(), Ls.data_ptr(), nbytes)

        def adj():
            e.adjoint_dev_ws(ca.ALS_SIDE_USERS, Y.data_ptr(), X.data_ptr(), Ls.data_ptr(), nbytes, g.data_ptr(),
                             xbar.data_ptr(), lbar.data_ptr(), ybar.data_ptr())

        def timed(fn):
            for _ in range(warmup):
                fn()
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            ev[0].record(s)
            for _ in range(steps):
                fn()
            ev[1].record(s)
            torch.cuda.synchronize()
            return ev[0].elapsed_time(ev[1]) / steps

        fwd_ms = timed(fwd)
        adj_ms = timed(adj)
        fails = int(e.info().failures)
    out = {"workload": f"ml-latest-shape {d.m}x{d.n}, {d.nnz} train ratings, f={f}, per-user lambda, fp16 Gram, "
                       "fp32 Cholesky; forward = users half-sweep keeping the factors, backward = adjoint",
           "forward_ms": fwd_ms, "adjoint_ms": adj_ms, "saved_factor_bytes": nbytes,
           "adjoint_saved_factor_read_tb_s": nbytes / (adj_ms * 1e-3) / 1e12, "failures": fails}
    log(f"[adjoint] {out}")
    return out


def batched_cg_leg(torch, E, steps, warmup):
    """The drop-in for the CG library's updateXWithCGHost (als_cg_update_x,
    and its FP16-A form): 16,384 synthetic SPD systems at f = 100 on
    caller buffers, 6 iterations from a zero start (no early exit at this
    tolerance), timed with HIP events; the bound is reading A once."""
    batch, f = 16384, 100
    g = torch.Generator(device="cuda").manual_seed(7)
    s = torch.cuda.Stream()
    out = {"workload": f"{batch} SPD systems, f={f}, A = M M^T / f + I (M ~ N(0,1)), cg_iters=6, tol=0",
           "a_bytes_fp32": batch * f * f * 4}
    with torch.cuda.stream(s):
        M = torch.randn(batch, f, f, device="cuda", generator=g)
        A = torch.bmm(M, M.transpose(1, 2).contiguous()) / f + torch.eye(f, device="cuda")
        del M
        b = torch.randn(batch, f, device="cuda", generator=g)
        for tag, AA in (("fp32_a", A), ("fp16_a", A.half())):
            x = torch.zeros(batch, f, device="cuda")
            for _ in range(warmup):
                x.zero_()
                E.cg_update_x(AA, x, b, 6, 0.0)
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            ms = 0.0
            for _ in range(steps):
                x.zero_()
                ev[0].record(s)
                E.cg_update_x(AA, x, b, 6, 0.0)
                ev[1].record(s)
                torch.cuda.synchronize()
                ms += ev[0].elapsed_time(ev[1])
            ms /= steps
            nbytes = AA.numel() * AA.element_size()
            out[tag] = {"ms": ms, "a_tb_s": nbytes / (ms * 1e-3) / 1e12}
    log(f"[batched cg] {out}")
    return out


def oracle_fixture(d, f):
    """The FP64 oracle's 10-sweep trajectory of this workload
    (tests/golden/netflix_trajectory.json), if it was made from the same
    generated arrays and f."""
    import hashlib
    p = ROOT / "tests" / "golden" / "netflix_trajectory.json"
    if not p.exists():
        return None
    fix = json.loads(p.read_text())
    if fix.get("f") != f:
        return None
    h = hashlib.sha256()
    for a in (d.csr_ptr, d.csr_idx, d.csr_val, d.csc_ptr, d.csc_idx, d.csc_val, d.t_row, d.t_col, d.t_val):
        h.update(a.tobytes())
    return fix if h.hexdigest() == fix["data_sha256"] else None


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(d, f, lam, budget_s=20.0):
    """FP64 oracle (Cholesky) on rows of both halves drawn uniformly with
    replacement, in doubling batches until the time budget or the row count
    is reached, scaled to a full sweep; timed with every host thread, then
    with one (a smaller sample). Returns the dict for the JSON line."""
    import numpy as np
    from oracle import als_oracle as O
    rng = np.random.default_rng(0)
    V = rng.random((d.n, f)) * 0.2
    U = rng.random((d.m, f)) * 0.2
    cfg = O.config(f, lam, O.SOLVE_CHOL)

    def sweep_seconds(budget):
        per_row = {}
        for name, rows, ptr, idx, val, Y in (("users", d.m, d.csr_ptr, d.csr_idx, d.csr_val, V),
                                              ("items", d.n, d.csc_ptr, d.csc_idx, d.csc_val, U)):
            k = 64
            spent, total_rows, total_t = 0.0, 0, 0.0
            while spent < budget / 2 and total_rows < rows:
                sel = rng.integers(0, rows, size=min(k, rows))
                t = time.perf_counter()
                O.update_rows(cfg, sel, ptr, idx, val, Y)
                el = time.perf_counter() - t
                total_t += el
                total_rows += sel.size
                spent += el
                k *= 2
            per_row[name] = (total_t / total_rows, total_rows)
        return per_row["users"][0] * d.m + per_row["items"][0] * d.n, per_row

    cores = int(O.lib.orc_num_threads())
    s_all, pr = sweep_seconds(budget_s)
    O.lib.orc_set_num_threads(1)
    s_one, pr1 = sweep_seconds(budget_s / 2)
    O.lib.orc_set_num_threads(cores)
    return {"value": 1.0 / s_all, "unit": "sweeps/s", "cores": cores, "kind": "port",
            "value_1core": 1.0 / s_one, "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": (f"FP64 oracle Cholesky update of {pr['users'][1]} user rows and {pr['items'][1]} item "
                       f"rows drawn uniformly with replacement ({cores} threads; {pr1['users'][1]} + "
                       f"{pr1['items'][1]} at 1 thread), per-row time scaled to all {d.m} users + {d.n} items")}


def copy_calibration(torch, nbytes: int, steps: int) -> int:
    """A plain device-to-device copy of known size: the HBM read/write counters
    of this kernel give the factor between counter units and bytes."""
    n = nbytes // 4
    src = torch.rand(n, device="cuda")
    dst = torch.empty_like(src)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        dst.copy_(src)
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record(s)
        for _ in range(steps):
            dst.copy_(src)
        t1.record(s)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / steps
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "ms": ms,
                      "GB_per_s": 2 * n * 4 / (ms * 1e-3) / 1e9}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--f", type=int, default=100)
    ap.add_argument("--prec", choices=["fp16", "bf16"], default="fp16")
    ap.add_argument("--refine", type=int, default=1)
    ap.add_argument("--refine-warm", type=int, default=1, choices=(0, 1),
                    help="1: the refinement starts from the previous iterate (residual formed before the "
                         "factorisation); the classic form is timed beside it unless --no-classic")
    ap.add_argument("--no-classic", action="store_true", help="skip timing the classic-refinement variant")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 comparison path")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-autotune", action="store_true", help="keep the cost-model schedule")
    ap.add_argument("--no-group", action="store_true", help="skip the three-model group leg (ml10M shape)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg (ml-latest shape)")
    ap.add_argument("--no-batched-cg", action="store_true", help="skip the batched-CG (caller buffers) leg")
    ap.add_argument("--scale", type=float, default=1.0, help="shrink the workload (debug only)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only time a plain device copy of BYTES (HBM counter calibration) and exit")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs on one GPU; multi-GPU row sharding is not built "
              "here. Run with --gpus 1.", file=sys.stderr)
        return 2

    import torch

    if a.copy_calibration:
        return copy_calibration(torch, a.copy_calibration, a.steps)

    import cuda_als_amd as ca
    from cuda_als_amd import engine as E

    t = time.perf_counter()
    wl = ca.netflix_shape() if a.scale == 1.0 else ca.scaled(ca.netflix_shape(), a.scale)
    d = wl.generate()
    log(f"generated {d.m} x {d.n}, {d.nnz} train / {d.n_test} test ratings in {time.perf_counter() - t:.1f}s")
    f, lam = a.f, 0.05
    prec = E.Prec.FP16 if a.prec == "fp16" else E.Prec.BF16
    fix = oracle_fixture(d, f) if a.scale == 1.0 else None
    log("FP64 oracle trajectory fixture:", "found" if fix else "not applicable")

    def run(cfg, tag, oname):
        r = time_path(torch, E, ca, d, cfg, a.steps, a.warmup, tag, not a.no_autotune,
                      (fix, oname) if fix else None)
        del r["eng"]
        return r

    mixed_cfg = E.make_config(f, lam, a.steps, E.SolveType.CHOL, prec, refine=a.refine,
                              refine_warm=bool(a.refine_warm), use_graph=True)
    M = run(mixed_cfg, "mixed", "chol")
    info_m = M["info"]
    classic = None
    if a.refine_warm and a.refine > 0 and not a.no_classic:
        c_cfg = E.make_config(f, lam, a.steps, E.SolveType.CHOL, prec, refine=a.refine, use_graph=True)
        C = run(c_cfg, "mixed-classic", "chol")
        classic = {"sweeps_per_s": a.steps / C["dt"], "users_ms": C["users_ms"], "items_ms": C["items_ms"],
                   "test_rmse": C["rmse"], "failures": int(C["info"].failures), "vs_oracle": C["vs_oracle"],
                   "solve": f"fp32 Cholesky + {a.refine} fp64-residual refinement step(s) of the fresh solve"}
    fp32 = cg6 = None
    if not a.no_fp32:
        fp32_cfg = E.make_config(f, lam, a.steps, E.SolveType.CHOL, E.Prec.FP32, use_graph=True)
        F = run(fp32_cfg, "fp32", "chol")
        fp32 = {"sweeps_per_s": a.steps / F["dt"], "users_ms": F["users_ms"], "items_ms": F["items_ms"],
                "test_rmse": F["rmse"], "failures": int(F["info"].failures), "vs_oracle": F["vs_oracle"],
                "solve": ("fp32 MFMA Gram, fp32 Cholesky, no refinement"
                          + (f"; rows under {list(F['info'].dual_limit)} ratings (users, items) through their "
                             f"n x n dual system" if F["info"].dual_max else ""))}
        cg_cfg = E.make_config(f, lam, a.steps, E.SolveType.CUMF_ALS_CG_FP32, E.Prec.FP32, cg_iters=6, cg_tol=1e-4,
                               use_graph=True)
        G = run(cg_cfg, "fp32-cg6", "cg6")
        cg6 = {"sweeps_per_s": a.steps / G["dt"], "users_ms": G["users_ms"], "items_ms": G["items_ms"],
               "test_rmse": G["rmse"], "failures": int(G["info"].failures), "vs_oracle": G["vs_oracle"],
               "solve": "fp32 MFMA Gram, cuMF CG recurrence: 6 iterations, warm start, exit on r.r < 1e-4 "
                        "(the reference's production solver)"}

    # Dominant kernel: the slower half-sweep of the mixed path (its update kernel plus the dual
    # kernel of its light rows, launched side by side from the engine's stream and timed together).
    um, im = M["users_ms"], M["items_ms"]
    half, ms = ("items", im) if im >= um else ("users", um)
    gram_flops = float(d.nnz) * f * (f + 1)           # lower triangle incl. diagonal, 2 flop per MAC
    achieved = gram_flops / (ms * 1e-3) / 1e12
    # The launch also factors one (f+1)x(f+1) system per row in FP32: (f+1)^3/3 MACs. Its time
    # bound at the FP32 matrix peak, added to the Gram's at the FP16 peak, gives the launch's
    # compute roofline time; frac_time compares it with the measured time.
    rows = d.m if half == "users" else d.n
    solve_flops = float(rows) * 2.0 * (f + 1) ** 3 / 3.0
    roof_ms = (gram_flops / (PEAK_FP16_TFLOPS * 1e12) + solve_flops / (PEAK_FP32_TFLOPS * 1e12)) * 1e3
    roofline = {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP16_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP16_TFLOPS, "traffic": parse_traffic(f"update_{half}"),
                "traffic_source": ("not measured in this run: HBM bytes per launch of the same kernels from "
                                   "profiles/traffic_update_netflix_shape.md (rocprofv3 --pmc FETCH_SIZE and "
                                   "WRITE_SIZE, scaled by a plain copy's counters)"),
                "kernel": (f"{half} half-sweep: update_kernel (16-bit Gram + Cholesky) and dual_kernel "
                           "(light rows), timed together with HIP events on the engine's stream"),
                "kernel_ms": ms, "flops_per_launch": gram_flops,
                "solve_fp32_flops_per_launch": solve_flops, "roofline_ms_gram_fp16_plus_solve_fp32": roof_ms,
                "frac_time": roof_ms / ms}
    # The other half beside it, against the memory roof: its algorithmic gather bytes (a 16-bit
    # factor row of the fixed side plus an index and a rating per rating) over its time.
    other, oms = ("users", um) if half == "items" else ("items", im)
    gather_bytes = float(d.nnz) * (2 * f + 8)
    gbs = gather_bytes / (oms * 1e-3) / 1e12
    roofline_other = {"half": other, "bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_TBS, "unit": "TB/s",
                      "frac": gbs / PEAK_HBM_TBS, "traffic": parse_traffic(f"update_{other}"),
                      "kernel_ms": oms, "bytes_per_launch": gather_bytes,
                      "note": "the fixed side's 16-bit rows largely come from the Infinity Cache and L2, so "
                              "traffic (L2 misses, from the committed profile) can sit below the algorithmic bytes"}

    line = {
        "metric": METRIC, "value": a.steps / M["dt"], "unit": "sweeps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": M["dt"] / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": a.prec, "data": "synthetic (seeded, seed 42)",
        "config": {"workload": f"netflix-shape {d.m}x{d.n}, {d.nnz} train ratings, f={f}, lambda={lam}",
                   "gram": f"{a.prec} MFMA, fp32 accumulate",
                   "solve": (f"fp32 Cholesky + {a.refine} fp64-residual refinement step(s)"
                             + (", the first from the previous iterate" if a.refine_warm and a.refine else "")
                             + (f"; rows under {list(info_m.dual_limit)} ratings (users, items; {info_m.dual_rows[0]} users, "
                                f"{info_m.dual_rows[1]} items) through their n x n dual system (fp32 "
                                f"Cholesky, no refinement step)"
                                if info_m.dual_max else "")),
                   "timing": "whole sweeps replayed from the captured sweep graph; per-half times from a "
                             "second pass launched half by half",
                   "users_ms": um, "items_ms": im, "test_rmse_after_steps": M["rmse"],
                   "failures": int(info_m.failures), "grid": list(info_m.grid), "lds_bytes": info_m.lds_bytes,
                   "schedule": M["tune"], "h2d_upload_s": M["upload_s"],
                   "sweeps_per_s_incl_h2d": a.steps / (M["dt"] + M["upload_s"])},
        "fp32_path": fp32,
        "fp32_cg6_path": cg6,
        "mixed_classic_refine": classic,
        "speedup_vs_fp32": (a.steps / M["dt"]) / fp32["sweeps_per_s"] if fp32 else None,
        "test_rmse_delta_vs_fp32": abs(M["rmse"] - fp32["test_rmse"]) if fp32 else None,
        "test_rmse_delta_vs_oracle": ({"mixed": M["vs_oracle"], "fp32": fp32 and fp32["vs_oracle"],
                                       "fp32_cg6": cg6 and cg6["vs_oracle"]} if fix else None),
        "roofline": roofline,
        "roofline_other_half": roofline_other,
        "multi_model_group": None if a.no_group else group_leg(torch, E, ca, a.steps, a.warmup,
                                                                not a.no_autotune),
        "differentiable_half_sweep": None if a.no_adjoint else adjoint_leg(torch, E, ca, a.steps, a.warmup),
        "batched_cg": None if a.no_batched_cg else batched_cg_leg(torch, E, a.steps, a.warmup),
        "cpu_baseline": None if a.no_cpu else cpu_baseline(d, f, lam),
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
