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


def test_argument_errors_return_status_without_device():
    """Entry points check their arguments before any device work: a NULL plan
    or a missing required buffer gives status -1 and a message, on any host."""
    from matscipy_neighbours_amd import _lib
    lib = _lib.load()
    z3 = np.zeros(3)
    eye = np.eye(3)
    pbc = np.ones(3, np.int32)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert lib.mnl_plan_lj(None, None, P(z3), P(eye), P(eye), P(pbc), 2, None, 1.0, 1.0, 1.0,
                           None, None, None) == -1
    assert b"plan is NULL" in lib.mnl_last_error()
    assert lib.mnl_plan_pair_backward(None, None, 0, 0, None, None, None, None, None, None,
                                      None, 0, None, 0, None, None) == -1
    assert b"plan is NULL" in lib.mnl_last_error()


@pytest.mark.skipif(HAS_GPU, reason="checks the cache query on a host without a device")
def test_result_cache_queries_without_device():
    """The kept-buffer query and release touch no device when nothing is
    kept: zero bytes for every device and for all, and a release is a no-op."""
    from matscipy_neighbours_amd import _lib
    lib = _lib.load()
    assert lib.mnl_cached_bytes(-1) == 0 and lib.mnl_cached_bytes(0) == 0
    lib.mnl_release_cached_memory()
    assert lib.mnl_cached_bytes(-1) == 0


def test_request_struct_layout_matches_header(tmp_path):
    """The ctypes mirror of struct mnl_request has the C layout: a small C
    program prints the size and field offsets the compiler gives."""
    import shutil
    import subprocess

    from matscipy_neighbours_amd import _lib
    cc = shutil.which("cc") or shutil.which("gcc")
    if cc is None:
        pytest.skip("no C compiler to read the header's layout")
    fields = [f[0] for f in _lib.Request._fields_]
    src = tmp_path / "layout.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"mnl.h\"\nint main(void){\n"
                   + 'printf("%zu\\n", sizeof(mnl_request));\n'
                   + "".join(f'printf("%zu\\n", offsetof(mnl_request, {f}));\n' for f in fields)
                   + "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)],
                   check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    assert got[0] == ctypes.sizeof(_lib.Request)
    assert got[1:] == [getattr(_lib.Request, f).offset for f in fields]


def test_request_and_order_arguments_without_device():
    """Argument errors of the request entry points and the order setter are
    reported without touching a device."""
    from matscipy_neighbours_amd import _lib
    lib = _lib.load()
    out = ctypes.c_void_p()
    assert lib.mnl_neighbour_list_request(None, ctypes.byref(out)) == -1
    assert b"request is NULL" in lib.mnl_last_error()
    assert lib.mnl_plan_set_cell_order(None, 1) == -1
    assert b"plan is NULL" in lib.mnl_last_error()
    assert lib.mnl_plan_cell_order(None) == -1
