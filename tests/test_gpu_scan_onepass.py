"""The single-pass scans (one launch with a decoupled look-back for scans of
2 to 1024 tiles of 2048 values) inside plan builds, against the oracle.

Their status words are cleared by the last tile of each scan, so the
sequence of builds on one plan matters: sizes of one, two, three and more
tiles in a row, the largest one-pass size (1024 tiles) and one tile more
(the reduce/spine/apply form), and a count + fill captured in a HIP graph
and replayed with moved atoms."""

import ctypes

import numpy as np
import pytest

import oracle
from conftest import assert_same_pairs

pytestmark = pytest.mark.gpu


def _uniform(n, density, seed):
    L = (n / density) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, L, (n, 3)), L * np.eye(3)


def _plan_build(lib, _lib, plan, pos_t, cell, rc):
    import torch
    n = int(pos_t.shape[0])
    inv = np.ascontiguousarray(np.linalg.inv(cell.T))
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    P_ = _lib.ptr
    npairs = ctypes.c_int64(0)
    _lib.check(lib.mnl_plan_count(plan.handle, sp, P_(np.zeros(3)), P_(np.ascontiguousarray(cell)),
                                  P_(inv), P_(np.ones(3, np.int32)), n, P_(pos_t), rc, None, None,
                                  0, None, -1, ctypes.byref(npairs)))
    P = int(npairs.value)
    out = {q: torch.empty(s, dtype=d, device=pos_t.device) for q, s, d in (
        ("i", (P,), torch.int32), ("j", (P,), torch.int32), ("D", (P, 3), torch.float64),
        ("d", (P,), torch.float64), ("S", (P, 3), torch.int32))}
    _lib.check(lib.mnl_plan_fill(plan.handle, sp, *(P_(out[q]) for q in "ijDdS")))
    counts = torch.empty(n, dtype=torch.int32, device=pos_t.device)
    _lib.check(lib.mnl_plan_counts(plan.handle, sp, P_(counts)))
    torch.cuda.synchronize()
    res = {q: t.cpu().numpy() for q, t in out.items()}
    res["c"] = counts.cpu().numpy()
    return res


def test_scan_sizes_in_sequence(gpu):
    """Builds of 1, 2, 3, 5, 2 and 1 tiles (of atoms) on one plan, each bit for
    bit against the oracle, counts included."""
    import torch

    from matscipy_neighbours_amd import _lib
    lib = _lib.load()
    plan = _lib.Plan(0)
    for k, n in enumerate((1500, 2500, 2048 * 2 + 7, 2048 * 4 + 100, 4096, 2048)):
        pos, cell = _uniform(n, 0.05, 100 + k)
        print(f"build {k}: {n} atoms", flush=True)
        res = _plan_build(lib, _lib, plan, torch.from_numpy(pos).cuda(), cell, 4.0)
        ref = oracle.neighbour_list("ijdDS", np.zeros(3), cell, np.linalg.inv(cell.T),
                                    np.ones(3, bool), pos, 4.0, with_counts=True)
        assert_same_pairs(res, ref, msg=f"build {k} ({n} atoms)")
        assert np.array_equal(res["c"], ref["counts"]), k


@pytest.mark.parametrize("tiles", [1024, 1025])
def test_largest_single_pass_scan(gpu, tiles):
    """2048 x 1024 atoms (the largest single-pass scan) and one tile more
    (the reduce/spine/apply form), sparse enough for few pairs: the rows
    follow the per-atom counts exactly."""
    import torch

    from matscipy_neighbours_amd import _lib
    lib = _lib.load()
    n = 2048 * tiles - 5
    pos, cell = _uniform(n, 0.01, tiles)
    print(f"{n} atoms", flush=True)
    plan = _lib.Plan(0)
    res = _plan_build(lib, _lib, plan, torch.from_numpy(pos).cuda(), cell, 2.5)
    i, c = res["i"], res["c"]
    ref_c = oracle.neighbour_list("i", np.zeros(3), cell, np.linalg.inv(cell.T),
                                  np.ones(3, bool), pos, 2.5, with_counts=True)["counts"]
    assert np.array_equal(c, ref_c)
    assert len(i) == int(ref_c.astype(np.int64).sum()) > 0
    assert np.array_equal(np.bincount(i, minlength=n), c)


def test_graph_replays_multi_tile(gpu):
    """count_async + fill_async of 20,000 atoms (ten tiles in the pair-count scan)
    captured in one graph and replayed three times with the atoms moved:
    every replay equals an eager build (the status words the scans clear
    after themselves are zero again for each replay)."""
    import torch

    from conftest import bits
    from matscipy_neighbours_amd import _lib, neighbour_list
    lib = _lib.load()
    n = 20_000
    pos, cell = _uniform(n, 0.05, 7)
    inv = np.ascontiguousarray(np.linalg.inv(cell.T))
    dev = torch.device("cuda:0")
    plan = _lib.Plan(0)
    P_ = _lib.ptr
    co, pb, ce = np.zeros(3), np.ones(3, np.int32), np.ascontiguousarray(cell)
    pos_t = torch.from_numpy(pos).to(dev)
    np_dev = torch.zeros(1, dtype=torch.int64, device=dev)
    sp = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(lib.mnl_plan_count_async(plan.handle, sp, P_(co), P_(ce), P_(inv), P_(pb), n,
                                        P_(pos_t), 4.0, None, None, 0, None, -1, P_(np_dev)))
    cap = 2 * int(np_dev.item()) + 1024
    b = {q: torch.empty(s, dtype=d, device=dev) for q, s, d in (
        ("i", (cap,), torch.int32), ("j", (cap,), torch.int32), ("D", (cap, 3), torch.float64),
        ("d", (cap,), torch.float64), ("S", (cap, 3), torch.int32))}
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        cs = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        _lib.check(lib.mnl_plan_count_async(plan.handle, cs, P_(co), P_(ce), P_(inv), P_(pb), n,
                                            P_(pos_t), 4.0, None, None, 0, None, -1,
                                            P_(np_dev)))
        _lib.check(lib.mnl_plan_fill_async(plan.handle, cs, cap, P_(b["i"]), P_(b["j"]),
                                           P_(b["D"]), P_(b["d"]), P_(b["S"]), 0))
    rng = np.random.default_rng(8)
    for rep in range(3):
        new = pos + rng.normal(0.0, 0.05, pos.shape)
        pos_t.copy_(torch.from_numpy(new))
        scratch = torch.randn(1 << 20, device=dev).sum()  # eager work in between
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        P = int(np_dev.item())
        assert 0 < P <= cap
        want = neighbour_list("ijDdS", positions=new, cell=cell, pbc=True, cutoff=4.0)
        for q, w in zip("ijDdS", want):
            assert np.array_equal(bits(b[q][:P].cpu().numpy()), bits(w)), f"replay {rep}: {q}"
        del scratch
