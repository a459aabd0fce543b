"""CPU-side checks of the C ABI: libmnl.so loads, exports every function that
include/*.h declares, and reports a clear error (no crash, no fallback) when
no GPU is present."""

import ctypes
import os
import re

import numpy as np
import pytest

from conftest import HAS_GPU, ROOT


def _declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(mnl_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = _declared_functions()
    for n in ("mnl_neighbour_list_device", "mnl_neighbour_count_device", "mnl_plan_count",
              "mnl_plan_fill", "mnl_result_export_dlpack", "mnl_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from matscipy_neighbours_amd import _lib
    lib = _lib.load()
    for name in sorted(_declared_functions()):
        assert hasattr(lib, name), f"{name} declared in include/ but not exported"


def test_binding_signatures_cover_header():
    from matscipy_neighbours_amd import _lib
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert bound == _declared_functions()


def test_version_string():
    from matscipy_neighbours_amd import _lib
    assert b"gfx950" in _lib.load().mnl_version()


@pytest.mark.skipif(HAS_GPU, reason="checks the no-GPU error path")
def test_no_gpu_fails_loudly():
    """Without a device the engine refuses to run; there is no CPU path."""
    from matscipy_neighbours_amd import _lib, neighbour_list
    lib = _lib.load()
    assert lib.mnl_device_count() == 0
    h = ctypes.c_void_p()
    assert lib.mnl_plan_create(-1, ctypes.byref(h)) == -1
    assert b"no HIP device" in lib.mnl_last_error()
    with pytest.raises(RuntimeError, match="no HIP device"):
        neighbour_list("ij", positions=np.zeros((2, 3)), cell=np.eye(3) * 5, pbc=True,
                       cutoff=1.0)
