"""Builds the engine library (gfx950 HIP + host C++) and the oracle.

    python cuda_als_amd/build.py          # everything
    python cuda_als_amd/build.py --lib    # engine only
(run by path: importing the package loads the library being rebuilt)

Outputs, in the source tree (git-ignored, shipped with the tree):
    cuda_als_amd/libals_mi355x.so   the product: kernels + C ABI (include/als.h,
                                    include/als_data.h)
    cuda_als_amd/als_main           reference-compatible driver (INI + .bin)
    oracle/libals_oracle.so         the FP64 checker (tests only)
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INC = ROOT / "include"
OBJ = PKG / "_build"
LIB = PKG / "libals_mi355x.so"
MAIN = PKG / "als_main"
ORACLE_SRC = ROOT / "oracle" / "als_oracle.c"
ORACLE_LIB = ROOT / "oracle" / "libals_oracle.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled for every HIP translation unit.
# MFMA accumulators are kept in VGPRs (not AGPRs): the same tiles feed the
# VALU solve right after the Gram, and one register pool lets two waves fit
# per SIMD at f = 100.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    f"-I{INC}", f"-I{CSRC}",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", f"-I{INC}", f"-I{CSRC}"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in r.stderr.splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{err}")


def build_lib() -> Path:
    OBJ.mkdir(exist_ok=True)
    jobs: list[list[str]] = []
    objs: list[Path] = []
    for p in (0, 1, 2):
        for unit in ("als_update",):
            o = OBJ / f"{unit}_p{p}.o"
            jobs.append([HIPCC, *HIP_FLAGS, f"-DALS_PREC={p}", "-c", str(CSRC / f"{unit}.hip"), "-o", str(o)])
            objs.append(o)
    # every other HIP unit of csrc/, plus the host units that call the HIP runtime
    hip_units = sorted(p.name for p in CSRC.glob("*.hip") if p.name != "als_update.hip")
    for src in (*hip_units, "als_engine.cpp", "als_sched.cpp"):
        o = OBJ / (Path(src).stem + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        jobs.append([HIPCC, *HIP_FLAGS, *lang, "-c", str(CSRC / src), "-o", str(o)])
        objs.append(o)
    for src in ("als_data.cpp", "als_model.cpp"):
        o = OBJ / (Path(src).stem + ".o")
        jobs.append(["g++", *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(o)])
        objs.append(o)
    workers = int(os.environ.get("MAX_JOBS", "8"))
    with cf.ThreadPoolExecutor(max_workers=min(workers, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lgomp"])
    os.replace(tmp, LIB)
    _run(["g++", *HOST_FLAGS, "-o", str(MAIN), str(CSRC / "als_main.cpp"), f"-L{PKG}", "-lals_mi355x",
          f"-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
    return LIB


def build_oracle() -> Path:
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run(["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-o", str(tmp), str(ORACLE_SRC), "-lm"])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true")
    ap.add_argument("--oracle", action="store_true")
    a = ap.parse_args(argv)
    both = not (a.lib or a.oracle)
    if a.oracle or both:
        print("built", build_oracle())
    if a.lib or both:
        print("built", build_lib())
    return 0


if __name__ == "__main__":
    sys.exit(main())
