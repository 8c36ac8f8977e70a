"""Shared test set-up: marker registration, import paths, fixture loaders."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "asfand-tmvp_amd"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; run with -m gpu")


def load_tv(name):
    """Arrays of a reference test-vector header, as written by make_golden.py."""
    doc = json.load(open(os.path.join(GOLDEN, name)))
    out = {}
    for key, rec in doc["arrays"].items():
        raw = bytes.fromhex(rec["hex"])
        out[key] = np.frombuffer(raw, "<u2").copy() if rec["ctype"] == "uint16_t" else raw
    return out


@pytest.fixture(scope="session")
def saber_tv():
    return load_tv("saber_tv.json")


@pytest.fixture(scope="session")
def sable_tv():
    return load_tv("sable_tv.json")
