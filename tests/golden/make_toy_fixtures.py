"""Writes the toy-data golden fixtures from the reference's data-prep logs.

The reference ships two tiny datasets whose preparation logs print every
array it wrote (data/toy_data/prepare_toy_data.log:62-75,
data/toy_data_2/prepare_toy_data_2.log:79-100): the train CSR and CSC and the
test COO, 0-based, in the reference's .bin layout. This script parses those
printed arrays (data, not code) into JSON fixtures:

    python tests/golden/make_toy_fixtures.py [/root/reference]

and records the run configuration the reference's configs give for them
(the key/value pairs of config/toy_data.ini and config/toy_data_2.ini: sizes,
f, lambda, sweeps, Gram type, solver, runs).
"""
from __future__ import annotations

import json
import re
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
KEYS = ["R_test_coo.data", "R_test_coo.row", "R_test_coo.col", "R_train_csr.data", "R_train_csr.indptr",
        "R_train_csr.indices", "R_train_csc.data", "R_train_csc.indptr", "R_train_csc.indices"]


def parse_log(text: str) -> dict:
    out = {}
    lines = text.splitlines()
    i = 0
    while i < len(lines):
        name = lines[i].strip()
        if name in KEYS and i + 1 < len(lines) and lines[i + 1].lstrip().startswith("["):
            buf = []
            j = i + 1
            while True:
                buf.append(lines[j])
                if "]" in lines[j]:
                    break
                j += 1
            nums = re.findall(r"-?\d+(?:\.\d*)?", " ".join(buf))
            out[name] = [float(x) if "data" in name else int(x) for x in nums]
            i = j + 1
            continue
        i += 1
    return out


def parse_ini(text: str) -> dict:
    """`section.key` -> value of an INI file (comments and blank lines dropped)."""
    out, section = {}, ""
    for line in text.splitlines():
        line = line.strip()
        if not line or line[0] in ";#":
            continue
        if line.startswith("["):
            section = line[1:line.index("]")].strip()
            continue
        if "=" in line:
            k, v = line.split("=", 1)
            out[f"{section}.{k.strip()}"] = v.strip()
    return out


def main(ref: str) -> None:
    root = Path(ref)
    specs = {
        "toy_data": (root / "data/toy_data/prepare_toy_data.log", root / "config/toy_data.ini", 15, 13),
        "toy_data_2": (root / "data/toy_data_2/prepare_toy_data_2.log", root / "config/toy_data_2.ini", 32, 23),
    }
    for name, (log, ini, m, n) in specs.items():
        arrays = parse_log(log.read_text())
        missing = [k for k in KEYS if k not in arrays]
        if missing:
            raise SystemExit(f"{log}: missing arrays {missing}")
        conf = parse_ini(ini.read_text())
        assert int(conf["data.m"]) == m and int(conf["data.n"]) == n, (ini, conf)
        # the reference's run configuration for this dataset (the values of its INI file)
        fixture = {"source": f"reference {log.relative_to(root)}", "m": m, "n": n, "arrays": arrays,
                   "ini_source": f"reference {ini.relative_to(root)}", "ini": conf}
        (HERE / f"{name}.json").write_text(json.dumps(fixture, indent=1))
        print("wrote", HERE / f"{name}.json")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
