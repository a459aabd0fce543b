"""The request entry point (the reference's NeighbourListRequest with its
CellOrder and device, plus a stream) and the plan's cell order, on the GPU.

Linear and Morton give the same pairs in the same order with the same bits
(both keep each cell's atoms in ascending index: the reference GPU path's
stable sort of the Morton keys, neighbour_list_gpu.cc:396-404, and its
Linear scatter): every golden case, both grid backends, against the
fixtures written by the compiled reference, through the request form with
host positions on the default stream and device positions on a side
stream, and through the plan API with the order set on the plan."""

import ctypes

import numpy as np
import pytest

from conftest import bits
from test_oracle import CASES, case_inputs, check_against_golden

pytestmark = pytest.mark.gpu

SHAPES = {"i": ((1,), np.int32), "j": ((1,), np.int32), "d": ((1,), np.float64),
          "D": ((1, 3), np.float64), "S": ((1, 3), np.int32)}


def _request(pos, cell, inv, pbc, cutoff, types, order, backend, on_device, stream):
    import torch

    from matscipy_neighbours_amd import _lib
    from matscipy_neighbours_amd import neighbours as nb
    lib = _lib.load()
    n = len(pos)
    numbers = types if types is not None else np.ones(n, np.int32)
    rc, per_atom, per_type_sq, nc = nb._parse_cutoff(cutoff, n)
    keep = []

    def host(a, dt):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return _lib.ptr(a)

    rq = _lib.Request()
    rq.quantities = 31
    rq.cell_origin = host(np.zeros(3), np.float64)
    rq.cell = host(cell, np.float64)
    rq.inv_cell = host(inv, np.float64)
    rq.pbc = host(np.asarray(pbc, bool), np.int32)
    rq.nat = n
    if on_device:
        dpos = torch.from_numpy(np.ascontiguousarray(pos)).to("cuda:0")
        torch.cuda.synchronize()
        keep.append(dpos)
        rq.positions = _lib.ptr(dpos)
    else:
        rq.positions = host(pos, np.float64)
    rq.positions_on_device = 1 if on_device else 0
    rq.cutoff = rc
    rq.per_atom_cutoff = host(per_atom, np.float64)
    rq.per_type_cutoff_sq = host(per_type_sq, np.float64)
    rq.ncutoffs = nc
    rq.types = host(numbers, np.int32) if per_type_sq is not None else None
    rq.order = order
    rq.device_id = 0
    rq.stream = stream
    rq.backend = backend
    res = ctypes.c_void_p()
    _lib.check(lib.mnl_neighbour_list_request(ctypes.byref(rq), ctypes.byref(res)))
    try:
        P = lib.mnl_result_npairs(res)
        out = {}
        for q, (shape, dt) in SHAPES.items():
            a = np.empty((P,) + shape[1:], dt)
            _lib.check(lib.mnl_result_copy_to_host(res, q.encode(), _lib.ptr(a)))
            out[q] = a
        return out
    finally:
        lib.mnl_result_free(res)


@pytest.mark.parametrize("name", CASES)
def test_linear_and_morton_identical_on_golden(gpu, name):
    import torch
    side = torch.cuda.Stream()
    pos, cell, inv, pbc, cutoff, types = case_inputs(name)
    for backend in (0, 1):
        lin = _request(pos, cell, inv, pbc, cutoff, types, 0, backend, False, None)
        check_against_golden(name, lin)
        mor = _request(pos, cell, inv, pbc, cutoff, types, 1, backend, True,
                       ctypes.c_void_p(side.cuda_stream))
        for q in "ijdDS":
            assert np.array_equal(bits(mor[q]), bits(lin[q])), (name, backend, q)


def test_plan_cell_order(gpu):
    import torch

    from matscipy_neighbours_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(12)
    cell = np.array([[14.0, 0.0, 0.0], [2.0, 13.0, 0.0], [0.5, 1.0, 12.0]])
    inv = np.ascontiguousarray(np.linalg.inv(cell.T))
    pos = torch.tensor(rng.uniform(0, 1, (4000, 3)) @ cell, device="cuda:0")
    pbc = np.ones(3, np.int32)
    co = np.zeros(3)
    plan = _lib.Plan(0)
    assert lib.mnl_plan_cell_order(plan.handle) == 0
    assert lib.mnl_plan_set_cell_order(plan.handle, 2) == -1
    assert b"unknown cell order" in lib.mnl_last_error()
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    got = []
    for order in (0, 1, 0):
        _lib.check(lib.mnl_plan_set_cell_order(plan.handle, order))
        assert lib.mnl_plan_cell_order(plan.handle) == order
        npairs = ctypes.c_int64()
        _lib.check(lib.mnl_plan_count(plan.handle, sp, _lib.ptr(co), _lib.ptr(cell),
                                      _lib.ptr(inv), _lib.ptr(pbc), len(pos), _lib.ptr(pos),
                                      3.0, None, None, 0, None, -1, ctypes.byref(npairs)))
        P = npairs.value
        o = [torch.empty(P, dtype=torch.int32, device="cuda:0"),
             torch.empty(P, dtype=torch.int32, device="cuda:0"),
             torch.empty((P, 3), dtype=torch.float64, device="cuda:0"),
             torch.empty(P, dtype=torch.float64, device="cuda:0"),
             torch.empty((P, 3), dtype=torch.int32, device="cuda:0")]
        _lib.check(lib.mnl_plan_fill(plan.handle, sp, *[_lib.ptr(t) for t in o]))
        torch.cuda.synchronize()
        got.append([t.cpu().numpy() for t in o])
    for a, b in zip(got[0], got[1]):
        assert np.array_equal(bits(a), bits(b))
    for a, b in zip(got[0], got[2]):
        assert np.array_equal(bits(a), bits(b))
