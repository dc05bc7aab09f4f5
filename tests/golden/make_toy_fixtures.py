"""Writes the toy-data golden fixtures from the reference's data-prep logs.

The reference ships two tiny datasets whose preparation logs print every
array it wrote (data/toy_data/prepare_toy_data.log:62-75,
data/toy_data_2/prepare_toy_data_2.log:79-100): the train CSR and CSC and the
test COO, 0-based, in the reference's .bin layout. This script parses those
printed arrays (data, not code) into JSON fixtures:

    python tests/golden/make_toy_fixtures.py [/root/reference]

and records the matrix sizes the reference's configs use for them
(config/toy_data.ini:9-12, config/toy_data_2.ini).
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


def main(ref: str) -> None:
    root = Path(ref)
    specs = {
        "toy_data": (root / "data/toy_data/prepare_toy_data.log", 15, 13),
        "toy_data_2": (root / "data/toy_data_2/prepare_toy_data_2.log", 32, 23),
    }
    for name, (log, m, n) in specs.items():
        arrays = parse_log(log.read_text())
        missing = [k for k in KEYS if k not in arrays]
        if missing:
            raise SystemExit(f"{log}: missing arrays {missing}")
        fixture = {"source": f"reference {log.relative_to(root)}", "m": m, "n": n, "arrays": arrays}
        (HERE / f"{name}.json").write_text(json.dumps(fixture, indent=1))
        print("wrote", HERE / f"{name}.json")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
