"""Shared test set-up: import paths for the product binding and the oracle,
and the `gpu` marker (tests that need an MI355X)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ALG_DIR = os.path.join(ROOT, "tests", "golden", "algorithms")
for p in (os.path.join(ROOT, "fmm-blis_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

ALG_NAMES = ["222", "232", "322", "333", "444"]


def alg_path(name):
    return os.path.join(ALG_DIR, f"{name}.txt")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
