"""The allocation-free entry point mnl_neighbour_list_noalloc, called through
the C ABI, against the plan API (mnl_plan_count + mnl_plan_fill) on the
same inputs, bit for bit: two calls in a row, the second with a smaller
input (the fixed device array keeps the first call's contents), then
capacity truncation, per-atom radii, the hashed grid, float32 output, the
size limit, and that a call leaves the device's free memory unchanged."""

import ctypes

import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu

Q = "ijDdS"


def _system(seed, n, edge, pbc, radii=False):
    rng = np.random.default_rng(seed)
    cell = np.diag([edge, edge * 1.1, edge * 0.9]) + rng.uniform(-0.5, 0.5, (3, 3))
    pos = rng.uniform(0, 1, (n, 3)) @ cell
    inv = np.ascontiguousarray(np.linalg.inv(cell.T))
    per_atom = rng.uniform(0.8, 1.6, n) if radii else None
    return dict(pos=pos, cell=np.ascontiguousarray(cell), inv=inv,
                pbc=np.asarray(pbc, np.int32), per_atom=per_atom)


def _dev(a, dtype=None):
    import torch
    return None if a is None else torch.tensor(np.ascontiguousarray(a), device="cuda:0",
                                               dtype=dtype)


def _plan_build(s, rc, backend, f32):
    """The existing entry point: plan count (waits) then fill."""
    import torch

    from matscipy_neighbours_amd import _lib
    lib = _lib.load()
    plan = _lib.Plan(0)
    pos = _dev(s["pos"], torch.float64)
    pa = _dev(s["per_atom"], torch.float64)
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    npairs = ctypes.c_int64(-1)
    co = np.zeros(3)
    _lib.check(lib.mnl_plan_count(plan.handle, sp, _lib.ptr(co), _lib.ptr(s["cell"]),
                                  _lib.ptr(s["inv"]), _lib.ptr(s["pbc"]), len(s["pos"]),
                                  _lib.ptr(pos), rc, _lib.ptr(pa), None, 0, None, backend,
                                  ctypes.byref(npairs)))
    P = npairs.value
    fdt = torch.float32 if f32 else torch.float64
    out = dict(i=torch.empty(P, dtype=torch.int32, device="cuda:0"),
               j=torch.empty(P, dtype=torch.int32, device="cuda:0"),
               D=torch.empty((P, 3), dtype=fdt, device="cuda:0"),
               d=torch.empty(P, dtype=fdt, device="cuda:0"),
               S=torch.empty((P, 3), dtype=torch.int32, device="cuda:0"))
    fill = lib.mnl_plan_fill_f32 if f32 else lib.mnl_plan_fill
    _lib.check(fill(plan.handle, sp, _lib.ptr(out["i"]), _lib.ptr(out["j"]), _lib.ptr(out["D"]),
                    _lib.ptr(out["d"]), _lib.ptr(out["S"])))
    torch.cuda.synchronize()
    return {q: t.cpu().numpy() for q, t in out.items()}


def _noalloc_build(s, rc, backend, f32, capacity, sentinel=-77):
    """The allocation-free entry point into buffers of `capacity` pairs plus
    a guard of 64 pairs filled with a sentinel."""
    import torch

    from matscipy_neighbours_amd import _lib
    lib = _lib.load()
    pos = _dev(s["pos"], torch.float64)
    pa = _dev(s["per_atom"], torch.float64)
    fdt = torch.float32 if f32 else torch.float64
    n = capacity + 64
    out = dict(i=torch.full((n,), sentinel, dtype=torch.int32, device="cuda:0"),
               j=torch.full((n,), sentinel, dtype=torch.int32, device="cuda:0"),
               D=torch.full((n, 3), sentinel, dtype=fdt, device="cuda:0"),
               d=torch.full((n,), sentinel, dtype=fdt, device="cuda:0"),
               S=torch.full((n, 3), sentinel, dtype=torch.int32, device="cuda:0"))
    total = torch.full((1,), -1, dtype=torch.int64, device="cuda:0")
    torch.cuda.synchronize()
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    co = np.zeros(3)
    print(f"noalloc: {len(s['pos'])} atoms, capacity {capacity}, backend {backend}, "
          f"f32 {f32}", flush=True)
    _lib.check(lib.mnl_neighbour_list_noalloc(
        sp, _lib.ptr(co), _lib.ptr(s["cell"]), _lib.ptr(s["inv"]), _lib.ptr(s["pbc"]),
        len(s["pos"]), _lib.ptr(pos), rc, _lib.ptr(pa), backend, capacity, _lib.ptr(out["i"]),
        _lib.ptr(out["j"]), _lib.ptr(out["D"]), _lib.ptr(out["d"]), _lib.ptr(out["S"]),
        1 if f32 else 0, _lib.ptr(total)))
    torch.cuda.synchronize()
    return {q: t.cpu().numpy() for q, t in out.items()}, int(total.item())


def _check(got, total, want, capacity, sentinel=-77):
    P = len(want["i"])
    assert total == P
    k = min(P, capacity)
    for q in Q:
        assert np.array_equal(bits(got[q][:k]), bits(want[q][:k])), q
        assert np.all(got[q][k:] == sentinel), q


def test_two_calls_in_a_row_second_smaller(gpu):
    big = _system(1, 20000, 30.0, (1, 1, 1))
    small = _system(2, 3000, 16.0, (1, 1, 0))
    want_big = _plan_build(big, 3.0, -1, False)
    got, total = _noalloc_build(big, 3.0, -1, False, len(want_big["i"]))
    _check(got, total, want_big, len(want_big["i"]))
    want_small = _plan_build(small, 2.5, -1, False)
    got, total = _noalloc_build(small, 2.5, -1, False, len(want_small["i"]) + 100)
    _check(got, total, want_small, len(want_small["i"]) + 100)


def test_capacity_radii_hashed_f32(gpu):
    s = _system(3, 5000, 20.0, (1, 0, 1), radii=True)
    for backend in (0, 1):
        for f32 in (False, True):
            want = _plan_build(s, 3.2, backend, f32)
            P = len(want["i"])
            got, total = _noalloc_build(s, 3.2, backend, f32, P)
            _check(got, total, want, P)
            got, total = _noalloc_build(s, 3.2, backend, f32, P // 3)
            _check(got, total, want, P // 3)


def test_limits_and_no_allocation(gpu):
    import torch

    from matscipy_neighbours_amd import _lib
    lib = _lib.load()
    s = _system(4, 8000, 24.0, (1, 1, 1))
    want = _plan_build(s, 3.0, -1, False)
    P = len(want["i"])
    pos = _dev(s["pos"], torch.float64)
    out = [torch.empty(P * 3, dtype=torch.float64, device="cuda:0") for _ in range(5)]
    total = torch.empty(1, dtype=torch.int64, device="cuda:0")
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    co = np.zeros(3)

    def call():
        return lib.mnl_neighbour_list_noalloc(
            sp, _lib.ptr(co), _lib.ptr(s["cell"]), _lib.ptr(s["inv"]), _lib.ptr(s["pbc"]),
            len(s["pos"]), _lib.ptr(pos), 3.0, None, -1, P, *[_lib.ptr(t) for t in out], 0,
            _lib.ptr(total))

    assert call() == 0
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(3):
        assert call() == 0
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    assert int(total.item()) == P
    # beyond the atom limit: refused on the host, nothing launched
    big = np.zeros((131073, 3))
    rc = lib.mnl_neighbour_list_noalloc(sp, _lib.ptr(co), _lib.ptr(s["cell"]),
                                        _lib.ptr(s["inv"]), _lib.ptr(s["pbc"]), len(big),
                                        _lib.ptr(pos), 3.0, None, -1, 0, None, None, None,
                                        None, None, 0, None)
    assert rc == -1 and b"at most 131072 atoms" in lib.mnl_last_error()
    # beyond the cell limit (a dense grid of 100^3 cells)
    huge = np.diag([300.0, 300.0, 300.0])
    hinv = np.ascontiguousarray(np.linalg.inv(huge.T))
    rc = lib.mnl_neighbour_list_noalloc(sp, _lib.ptr(co), _lib.ptr(huge), _lib.ptr(hinv),
                                        _lib.ptr(s["pbc"]), 100, _lib.ptr(pos), 3.0, None, 0,
                                        0, None, None, None, None, None, 0, None)
    assert rc == -1 and b"allocation-free workspace" in lib.mnl_last_error()
