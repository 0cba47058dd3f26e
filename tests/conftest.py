"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on the CPU."""
import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libnli.so kernels)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    return O.Oracle()


@pytest.fixture(scope="session")
def reference():
    import oracle as O
    if not os.path.exists(O.REF_SO):
        pytest.skip("reference core not built in place (oracle/_ref, needs the reference sources)")
    return O.Reference()


@pytest.fixture(scope="session")
def nli():
    return importlib.import_module("non-local-inpainting_amd")


@pytest.fixture(scope="session")
def syn():
    return importlib.import_module("non-local-inpainting_amd.synthetic")


@pytest.fixture(scope="session")
def ctx(nli):
    c = nli.Context(0)
    yield c
    c.close()


def golden(name):
    path = os.path.join(GOLDEN, name + ".npz")
    if not os.path.exists(path):
        pytest.skip(f"missing fixture {path}")
    return np.load(path, allow_pickle=False)


def bitexact(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return False
    if a.dtype == np.float32:
        return np.array_equal(a.view(np.uint32), b.view(np.uint32))
    if a.dtype == np.float64:
        return np.array_equal(a.view(np.uint64), b.view(np.uint64))
    return np.array_equal(a, b)
