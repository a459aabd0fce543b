"""ctypes binding of the mnl C ABI (include/mnl.h).

The engine is libmnl.so next to this file (built by
matscipy-neighbours_amd/csrc/Makefile). There is no CPU fallback: importing
works without a GPU, but every compute call raises RuntimeError when the
library is missing or no HIP device is present.
"""

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmnl.so")

QUANTITY_BITS = {"i": 1, "j": 2, "D": 4, "d": 8, "S": 16}
BACKEND_AUTO, BACKEND_DENSE, BACKEND_HASHED = -1, 0, 1
ORDER_LINEAR, ORDER_MORTON = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lp = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p



class Request(ctypes.Structure):
    """struct mnl_request: the reference's NeighbourListRequest
    (src/libneighbours/neighbour_list_gpu.hh:49-65) plus stream and backend."""
    _fields_ = [("quantities", ctypes.c_int), ("cell_origin", _vp), ("cell", _vp),
                ("inv_cell", _vp), ("pbc", _vp), ("nat", ctypes.c_int32),
                ("positions", _vp), ("positions_on_device", ctypes.c_int),
                ("cutoff", ctypes.c_double), ("per_atom_cutoff", _vp),
                ("per_type_cutoff_sq", _vp), ("ncutoffs", ctypes.c_int32), ("types", _vp),
                ("order", ctypes.c_int), ("device_id", ctypes.c_int), ("stream", _vp),
                ("backend", ctypes.c_int)]


# (name, restype, argtypes) of every entry point declared in include/mnl.h.
SIGNATURES = [
    ("mnl_last_error", ctypes.c_char_p, []),
    ("mnl_version", ctypes.c_char_p, []),
    ("mnl_device_count", ctypes.c_int, []),
    ("mnl_neighbour_list_device", ctypes.c_int,
     [ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_int32, _vp, ctypes.c_int,
      ctypes.c_double, _vp, _vp, ctypes.c_int32, _vp, ctypes.c_int,
      ctypes.POINTER(_vp)]),
    ("mnl_neighbour_count_device", ctypes.c_int,
     [_vp, _vp, _vp, _vp, ctypes.c_int32, _vp, ctypes.c_int, ctypes.c_double,
      _vp, _vp, ctypes.c_int32, _vp, ctypes.c_int, ctypes.POINTER(_vp)]),
    ("mnl_neighbour_list_request", ctypes.c_int,
     [ctypes.POINTER(Request), ctypes.POINTER(_vp)]),
    ("mnl_neighbour_count_request", ctypes.c_int,
     [ctypes.POINTER(Request), ctypes.POINTER(_vp)]),
    ("mnl_plan_set_cell_order", ctypes.c_int, [_vp, ctypes.c_int]),
    ("mnl_plan_cell_order", ctypes.c_int, [_vp]),
    ("mnl_result_npairs", ctypes.c_int64, [_vp]),
    ("mnl_result_nat", ctypes.c_int32, [_vp]),
    ("mnl_result_device", ctypes.c_int, [_vp]),
    ("mnl_result_data", _vp, [_vp, ctypes.c_char]),
    ("mnl_result_copy_to_host", ctypes.c_int, [_vp, ctypes.c_char, _vp]),
    ("mnl_result_export_dlpack", _vp, [_vp, ctypes.c_char]),
    ("mnl_result_free", None, [_vp]),
    ("mnl_release_cached_memory", None, []),
    ("mnl_cached_bytes", ctypes.c_int64, [ctypes.c_int]),
    ("mnl_plan_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    ("mnl_plan_destroy", None, [_vp]),
    ("mnl_plan_device", ctypes.c_int, [_vp]),
    ("mnl_plan_count", ctypes.c_int,
     [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int32, _vp, ctypes.c_double, _vp,
      _vp, ctypes.c_int32, _vp, ctypes.c_int, _lp]),
    ("mnl_plan_count_async", ctypes.c_int,
     [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int32, _vp, ctypes.c_double, _vp,
      _vp, ctypes.c_int32, _vp, ctypes.c_int, _vp]),
    ("mnl_plan_fill_async", ctypes.c_int,
     [_vp, _vp, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, ctypes.c_int]),
    ("mnl_plan_fill", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("mnl_neighbour_list_noalloc", ctypes.c_int,
     [_vp, _vp, _vp, _vp, _vp, ctypes.c_int32, _vp, ctypes.c_double, _vp, ctypes.c_int,
      ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    ("mnl_plan_fill_f32", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("mnl_plan_counts", ctypes.c_int, [_vp, _vp, _vp]),
    ("mnl_plan_offsets", _vp, [_vp]),
    ("mnl_plan_backend", ctypes.c_int, [_vp]),
    ("mnl_plan_matrix", ctypes.c_int,
     [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int32, _vp, ctypes.c_double, _vp, _vp,
      ctypes.c_int32, _vp, ctypes.c_int, ctypes.c_int32, _vp, _vp, _vp,
      ctypes.POINTER(ctypes.c_int32)]),
    ("mnl_plan_matrix_async", ctypes.c_int,
     [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int32, _vp, ctypes.c_double, _vp, _vp,
      ctypes.c_int32, _vp, ctypes.c_int, ctypes.c_int32, _vp, _vp, _vp, _vp]),
    ("mnl_plan_set_batch", ctypes.c_int,
     [_vp, _vp, ctypes.c_int32, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("mnl_plan_batch_count", ctypes.c_int, [_vp, _vp, _vp, _lp]),
    ("mnl_plan_batch_count_async", ctypes.c_int, [_vp, _vp, _vp, _vp]),
    ("mnl_plan_system_offsets", ctypes.c_int, [_vp, _vp, _vp]),
    ("mnl_plan_nsys", ctypes.c_int32, [_vp]),
    ("mnl_plan_batch_graph_create", ctypes.c_int,
     [_vp, _vp, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp,
      ctypes.POINTER(_vp)]),
    ("mnl_graph_launch", ctypes.c_int, [_vp, _vp]),
    ("mnl_graph_destroy", None, [_vp]),
    ("mnl_plan_pair_backward", ctypes.c_int,
     [_vp, _vp, ctypes.c_int64, ctypes.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
      ctypes.c_int, _vp, ctypes.c_int, _vp, _vp]),
    ("mnl_plan_autotune", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("mnl_autotune_clear", None, []),
    ("mnl_plan_set_launch_shape", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32]),
    ("mnl_plan_lj", ctypes.c_int,
     [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int32, _vp, ctypes.c_double, ctypes.c_double,
      ctypes.c_double, _vp, _vp, _vp]),
    ("mnl_first_neighbours", ctypes.c_int,
     [_vp, ctypes.c_int32, ctypes.c_int64, _vp, _vp]),
    ("mnl_plan_count_distinct", ctypes.c_int,
     [_vp, _vp, ctypes.c_int64, _vp, ctypes.POINTER(ctypes.c_int32)]),
    ("mnl_plan_triplets_count", ctypes.c_int,
     [_vp, _vp, ctypes.c_int32, _vp, ctypes.c_int64, _vp, ctypes.c_double, _lp]),
    ("mnl_plan_triplets_fill", ctypes.c_int, [_vp, _vp, _vp, _vp]),
    ("mnl_debug_sqrt", ctypes.c_int, [_vp, _vp, ctypes.c_int64]),
]

_lib = None
_lock = threading.Lock()


def load():
    """Load libmnl.so (raises RuntimeError if it was not built).

    PyTorch-ROCm ships its own HIP runtime (libamdhip64.so, SONAME
    libamdhip64.so.7). torch is imported first so that libmnl.so's
    libamdhip64.so.7 dependency binds to that already-loaded runtime: one HIP
    runtime per process, so torch tensors, streams and the engine share
    device pointers and queues. Outside Python the library uses the system
    ROCm runtime."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"mnl engine library not built: {LIB_PATH}. Build it with "
                    "`make -C matscipy-neighbours_amd/csrc` (or "
                    "__graft_entry__.build()).")
            try:
                import torch  # noqa: F401  (binds the shared HIP runtime)
            except ImportError:  # pragma: no cover - torch-less use of the C ABI
                pass
            lib = ctypes.CDLL(LIB_PATH)
            for name, res, args in SIGNATURES:
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def last_error():
    return load().mnl_last_error().decode(errors="replace")


def check(status):
    if status != 0:
        raise RuntimeError(last_error())


def device_count():
    return int(load().mnl_device_count())


def require_gpu():
    """Fail loudly when no HIP device is usable (there is no CPU fallback)."""
    if device_count() <= 0:
        raise RuntimeError("mnl: no HIP device available; the neighbour-list "
                           "engine runs only on the GPU")


def ptr(x):
    """Raw address of a numpy array or torch tensor as an int (None -> NULL);
    the bound entry points declare their pointers c_void_p, which takes it
    as is."""
    if x is None:
        return None
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    return x.__array_interface__["data"][0]


class Plan:
    """Owning wrapper of an mnl_plan (reusable device workspace)."""

    def __init__(self, device):
        lib = load()
        h = ctypes.c_void_p()
        check(lib.mnl_plan_create(int(device), ctypes.byref(h)))
        self._h = h
        self.device = int(device)
        self.lock = threading.Lock()

    @property
    def handle(self):
        return self._h

    def __del__(self):
        try:
            if self._h:
                load().mnl_plan_destroy(self._h)
                self._h = None
        except Exception:
            pass


_plans = {}


def plan_for(device):
    """One cached plan per device."""
    with _lock:
        p = _plans.get(device)
    if p is None:
        p = Plan(device)
        with _lock:
            _plans.setdefault(device, p)
            p = _plans[device]
    return p
