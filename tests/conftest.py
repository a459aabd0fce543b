"""Test configuration: import paths, the `gpu` marker and shared helpers."""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "matscipy-neighbours_amd")
for p in (ROOT, PKG_DIR):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def _has_gpu():
    try:
        from matscipy_neighbours_amd import _lib
        return _lib.device_count() > 0
    except Exception:
        return False


HAS_GPU = _has_gpu()


@pytest.fixture(scope="session")
def gpu():
    if not HAS_GPU:
        pytest.fail("GPU test collected but no HIP device / libmnl.so is available")
    return True


def bits(a):
    """Bit pattern of an array, for exact comparisons of floats."""
    a = np.ascontiguousarray(a)
    if a.dtype == np.float64:
        return a.view(np.uint64)
    if a.dtype == np.float32:
        return a.view(np.uint32)
    return a


def assert_same_pairs(got, ref, quantities="ijdDS", msg=""):
    """Bit-exact equality of every requested quantity (pairs, order, bits)."""
    for q in quantities:
        g = np.asarray(got[q])
        r = np.asarray(ref[q])
        assert g.shape == r.shape, f"{msg} {q}: shape {g.shape} vs {r.shape}"
        if not np.array_equal(bits(g), bits(r)):
            bad = np.nonzero((bits(g) != bits(r)).reshape(len(g), -1).any(axis=1))[0]
            raise AssertionError(f"{msg} {q}: {len(bad)} rows differ, first at {bad[:5]}")
