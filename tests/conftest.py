"""Shared fixtures: the product binding (libgpb.so via ctypes), the CPU oracle
(liboracle.so, test infrastructure) and a GPU context for `gpu` tests."""
import importlib.util
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _load(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def gpb():
    return _load("gpb_amd_binding", ROOT / "ucm-mpi-gpu_amd" / "gpb.py")


@pytest.fixture(scope="session")
def orc():
    return _load("gpb_oracle_binding", ROOT / "oracle" / "oracle.py")


@pytest.fixture(scope="session")
def ctx(gpb):
    if gpb.device_count() < 1:
        pytest.fail("no HIP device visible: gpu tests must run on a GPU box")
    c = gpb.Context(0)
    yield c
    c.close()


def synth_packed(gpb, w, h, regions=20, sigma=6.0, seed=1, classes=1):
    rgb = gpb.synth_voronoi(w, h, regions, sigma, seed, classes)
    return rgb, gpb.pack_rgb(rgb)
