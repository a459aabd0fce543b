"""The plan API's asynchronous forms through the C ABI, against the oracle or
the synchronous forms: count and fill with a caller-chosen capacity
(truncation leaves the rest of the buffers untouched), the backend and
offsets accessors, a single-system count + fill captured in a HIP graph
(torch.cuda.graph) and replayed with moved atoms, the matrix's device-side
overflow flag, and the asynchronous batch count."""

import ctypes

import numpy as np
import pytest

import oracle
from conftest import bits

pytestmark = pytest.mark.gpu


def _geom(cell, pbc):
    cell = np.ascontiguousarray(cell, dtype=float)
    inv = np.ascontiguousarray(np.linalg.inv(cell.T))
    pb = np.ascontiguousarray(np.broadcast_to(np.asarray(pbc, bool), 3), dtype=np.int32)
    return np.zeros(3), cell, inv, pb


def _bufs(torch, n, dev, fill=True):
    b = dict(i=torch.empty(n, dtype=torch.int32, device=dev),
             j=torch.empty(n, dtype=torch.int32, device=dev),
             D=torch.empty((n, 3), dtype=torch.float64, device=dev),
             d=torch.empty(n, dtype=torch.float64, device=dev),
             S=torch.empty((n, 3), dtype=torch.int32, device=dev))
    if fill:
        for t in b.values():
            t.fill_(-7)
    return b


def _system(seed, n=3000, rc=2.0):
    rng = np.random.default_rng(seed)
    cell = np.diag([14.0, 15.0, 16.0])
    cell[1, 0] = 1.5
    pos = rng.uniform(0.0, 1.0, (n, 3)) @ cell
    return pos, cell, rc


def test_async_count_fill_capacity(gpu):
    import torch

    from matscipy_neighbours_amd import _lib
    lib = _lib.load()
    pos, cell, rc = _system(91)
    n = len(pos)
    co, ce, inv, pb = _geom(cell, True)
    ref = oracle.neighbour_list("ijdDS", co, ce, inv, np.ones(3, bool), pos, rc, with_counts=True)
    P_ref = len(ref["i"])
    dev = torch.device("cuda:0")
    plan = _lib.Plan(0)
    sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    P_ = _lib.ptr
    pos_t = torch.from_numpy(pos).to(dev)
    np_dev = torch.zeros(1, dtype=torch.int64, device=dev)
    for backend in (0, 1):
        _lib.check(lib.mnl_plan_count_async(plan.handle, sp, P_(co), P_(ce), P_(inv), P_(pb), n,
                                            P_(pos_t), rc, None, None, 0, None, backend,
                                            P_(np_dev)))
        assert int(np_dev.item()) == P_ref
        assert lib.mnl_plan_backend(plan.handle) == backend
        assert lib.mnl_plan_offsets(plan.handle)
        sys_off = torch.empty(2, dtype=torch.int64, device=dev)
        _lib.check(lib.mnl_plan_system_offsets(plan.handle, sp, P_(sys_off)))
        assert sys_off.cpu().tolist() == [0, P_ref]
        for cap in (P_ref, P_ref // 2, 0):
            b = _bufs(torch, P_ref, dev)
            _lib.check(lib.mnl_plan_fill_async(plan.handle, sp, cap, P_(b["i"]), P_(b["j"]),
                                               P_(b["D"]), P_(b["d"]), P_(b["S"]), 0))
            torch.cuda.synchronize()
            for q in "ijDdS":
                got = b[q].cpu().numpy()
                assert np.array_equal(bits(got[:cap]), bits(ref[q][:cap])), (backend, cap, q)
                # beyond the capacity nothing is written
                rest = got[cap:]
                sentinel = np.full_like(rest, -7)
                assert np.array_equal(bits(rest), bits(sentinel)), (backend, cap, q)


def test_plan_offsets_pointer(gpu):
    """mnl_plan_offsets: the int64 [nat+1] CSR offsets of the i-sorted list,
    read through __cuda_array_interface__ where torch supports it."""
    import torch

    from matscipy_neighbours_amd import _lib
    lib = _lib.load()
    pos, cell, rc = _system(92, n=1500)
    n = len(pos)
    co, ce, inv, pb = _geom(cell, True)
    ref = oracle.neighbour_list("ij", co, ce, inv, np.ones(3, bool), pos, rc, with_counts=True)
    plan = _lib.Plan(0)
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    pos_t = torch.from_numpy(pos).cuda()
    npairs = ctypes.c_int64(0)
    _lib.check(lib.mnl_plan_count(plan.handle, sp, _lib.ptr(co), _lib.ptr(ce), _lib.ptr(inv),
                                  _lib.ptr(pb), n, _lib.ptr(pos_t), rc, None, None, 0, None, -1,
                                  ctypes.byref(npairs)))
    ptr = lib.mnl_plan_offsets(plan.handle)

    class _View:
        __cuda_array_interface__ = {"shape": (n + 1,), "typestr": "<i8", "data": (ptr, False),
                                    "version": 3, "strides": None}
    try:
        offs = torch.as_tensor(_View(), device="cuda")
    except (TypeError, RuntimeError, ValueError) as e:
        pytest.skip(f"this torch build cannot import __cuda_array_interface__: {e}")
    want = np.concatenate([[0], np.cumsum(ref["counts"])])
    assert np.array_equal(offs.cpu().numpy(), want)


def test_async_build_captured_in_graph(gpu):
    """count_async + fill_async captured in a torch CUDA graph (after one
    eager build sized the workspaces), replayed three times with the atoms
    moved in place and eager work and synchronisations in between: every
    replay equals an eager build of the same positions."""
    import torch

    from matscipy_neighbours_amd import _lib, neighbour_list
    lib = _lib.load()
    pos, cell, rc = _system(93, n=2500)
    n = len(pos)
    co, ce, inv, pb = _geom(cell, True)
    dev = torch.device("cuda:0")
    plan = _lib.Plan(0)
    P_ = _lib.ptr
    pos_t = torch.from_numpy(pos).to(dev)
    np_dev = torch.zeros(1, dtype=torch.int64, device=dev)
    sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(lib.mnl_plan_count_async(plan.handle, sp, P_(co), P_(ce), P_(inv), P_(pb), n,
                                        P_(pos_t), rc, None, None, 0, None, -1, P_(np_dev)))
    cap = 2 * int(np_dev.item()) + 1024
    b = _bufs(torch, cap, dev, fill=False)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        cs = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        _lib.check(lib.mnl_plan_count_async(plan.handle, cs, P_(co), P_(ce), P_(inv), P_(pb), n,
                                            P_(pos_t), rc, None, None, 0, None, -1,
                                            P_(np_dev)))
        _lib.check(lib.mnl_plan_fill_async(plan.handle, cs, cap, P_(b["i"]), P_(b["j"]),
                                           P_(b["D"]), P_(b["d"]), P_(b["S"]), 0))
    rng = np.random.default_rng(94)
    for rep in range(3):
        new = pos + rng.normal(0.0, 0.05, pos.shape)
        pos_t.copy_(torch.from_numpy(new))
        scratch = torch.randn(1 << 20, device=dev).sum()  # eager work in between
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        P = int(np_dev.item())
        assert 0 < P <= cap
        want = neighbour_list("ijDdS", positions=new, cell=cell, pbc=True, cutoff=rc)
        for q, w in zip("ijDdS", want):
            assert np.array_equal(bits(b[q][:P].cpu().numpy()), bits(w)), f"replay {rep}: {q}"
        del scratch


def test_matrix_async_matches_sync(gpu):
    """mnl_plan_matrix_async with its device overflow flag against the
    waiting form, with a capacity that fits and one that does not."""
    import torch

    from matscipy_neighbours_amd import _lib
    lib = _lib.load()
    pos, cell, rc = _system(95, n=1200)
    n = len(pos)
    co, ce, inv, pb = _geom(cell, True)
    dev = torch.device("cuda:0")
    plan = _lib.Plan(0)
    sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    P_ = _lib.ptr
    pos_t = torch.from_numpy(pos).to(dev)
    for K in (48, 3):
        out = []
        for mode in ("sync", "async"):
            idx = torch.full((n, K), -7, dtype=torch.int32, device=dev)
            dist = torch.full((n, K, 3), -7.0, dtype=torch.float64, device=dev)
            cnt = torch.full((n,), -7, dtype=torch.int32, device=dev)
            if mode == "sync":
                ov = ctypes.c_int32(-1)
                _lib.check(lib.mnl_plan_matrix(plan.handle, sp, P_(co), P_(ce), P_(inv), P_(pb), n,
                                               P_(pos_t), rc, None, None, 0, None, -1, K,
                                               P_(idx), P_(dist), P_(cnt), ctypes.byref(ov)))
                flag = ov.value
            else:
                ovd = torch.full((1,), -1, dtype=torch.int32, device=dev)
                _lib.check(lib.mnl_plan_matrix_async(plan.handle, sp, P_(co), P_(ce), P_(inv),
                                                     P_(pb), n, P_(pos_t), rc, None, None, 0,
                                                     None, -1, K, P_(idx), P_(dist), P_(cnt),
                                                     P_(ovd)))
                torch.cuda.synchronize()
                flag = int(ovd.item())
            out.append((idx.cpu().numpy(), dist.cpu().numpy(), cnt.cpu().numpy(), flag))
        (i0, d0, c0, f0), (i1, d1, c1, f1) = out
        assert np.array_equal(i0, i1) and np.array_equal(bits(d0), bits(d1))
        assert np.array_equal(c0, c1)
        assert f0 == f1 == (1 if c0.max() > K else 0)


def test_batch_count_async(gpu):
    """mnl_plan_batch_count_async + mnl_plan_fill_async + the per-system
    offsets against neighbour_list_batch (the waiting batch path)."""
    import torch

    from matscipy_neighbours_amd import _lib, neighbour_list_batch, workloads
    from matscipy_neighbours_amd.neighbours import _batch_geometry
    lib = _lib.load()
    systems = workloads.uneven_batch(nsys=96, seed=31)
    pos = np.ascontiguousarray(np.concatenate([s["positions"] for s in systems]))
    cells = np.stack([s["cell"] for s in systems])
    pbc = np.stack([np.asarray(s["pbc"], bool) for s in systems])
    rc = np.array([s["cutoff"] for s in systems], dtype=float)
    nat = np.array([len(s["positions"]) for s in systems])
    nsys, offs, co, ce, inv, pb, rcs = _batch_geometry(cells, pbc, rc, nat, None)
    dev = torch.device("cuda:0")
    plan = _lib.Plan(0)
    sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    P_ = _lib.ptr
    pos_t = torch.from_numpy(pos).to(dev)
    _lib.check(lib.mnl_plan_set_batch(plan.handle, sp, nsys, P_(offs), P_(co), P_(ce), P_(inv),
                                      P_(pb), P_(rcs)))
    assert lib.mnl_plan_nsys(plan.handle) == nsys
    np_dev = torch.zeros(1, dtype=torch.int64, device=dev)
    _lib.check(lib.mnl_plan_batch_count_async(plan.handle, sp, P_(pos_t), P_(np_dev)))
    P = int(np_dev.item())
    b = _bufs(torch, P, dev)
    _lib.check(lib.mnl_plan_fill_async(plan.handle, sp, P, P_(b["i"]), P_(b["j"]), P_(b["D"]),
                                       P_(b["d"]), P_(b["S"]), 0))
    sys_off = torch.empty(nsys + 1, dtype=torch.int64, device=dev)
    _lib.check(lib.mnl_plan_system_offsets(plan.handle, sp, P_(sys_off)))
    torch.cuda.synchronize()
    want = neighbour_list_batch("ijDdS", pos, cells, pbc, rc, nat)
    got = [b[q].cpu().numpy() for q in "ijDdS"] + [sys_off.cpu().numpy()]
    for q, a, w in zip("ijDdS+", got, want):
        assert np.array_equal(bits(a), bits(w)), q
