"""Test configuration: the `gpu` marker and shared fixtures.

GPU tests call the engine through its C ABI (cuda_als_amd/libals_mi355x.so)
and compare with the FP64 oracle (oracle/libals_oracle.so) on the same
seeded inputs. CPU tests cover the oracle, the host logic and the library's
exported symbols.
"""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def toy():
    import cuda_als_amd as ca
    return ca.toy_1k().generate()


GOLDEN = Path(__file__).resolve().parent / "golden"


def reference_toy(name):
    """The reference's toy dataset `name` (its prep-log arrays) as RatingData,
    and the values of its INI file (`section.key` -> str)."""
    import json

    import cuda_als_amd as ca
    d = json.loads((GOLDEN / f"{name}.json").read_text())
    a = d["arrays"]
    data = ca.from_arrays(d["m"], d["n"], a["R_train_csr.indptr"], a["R_train_csr.indices"], a["R_train_csr.data"],
                          a["R_test_coo.row"], a["R_test_coo.col"], a["R_test_coo.data"])
    return data, d["ini"]
