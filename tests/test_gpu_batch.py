"""Batched builds: many independent systems in one build must give, for every
system, exactly the pairs (order and bits) of a single-system build of it,
which the oracle (the reference CPU path restated) fixes.

Covers the uneven 1024-system workload (8-512 atoms, tilted cubic cells at
0.05 atoms/A^3, rc 5 A: boxes narrower than twice the cutoff, where one atom
is a neighbour through several images), empty systems, mixed periodicity and per-system cutoffs, float32
outputs, and the captured HIP graph replayed several times with eager work
in between (every replay checked against the eager build)."""

import ctypes

import numpy as np
import pytest

import oracle
from conftest import bits

pytestmark = pytest.mark.gpu


def _concat(systems):
    pos = np.concatenate([s["positions"] for s in systems]) if systems else np.zeros((0, 3))
    cells = np.stack([s["cell"] for s in systems])
    pbc = np.stack([np.asarray(s["pbc"], bool) for s in systems])
    rc = np.array([s["cutoff"] for s in systems], dtype=float)
    nat = np.array([len(s["positions"]) for s in systems])
    return np.ascontiguousarray(pos), cells, pbc, rc, nat


def _oracle_system(s):
    cell = np.ascontiguousarray(s["cell"], dtype=float)
    inv = np.ascontiguousarray(np.linalg.inv(cell.T))
    return oracle.neighbour_list("ijdDS", np.zeros(3), cell, inv, np.asarray(s["pbc"], bool),
                                 np.ascontiguousarray(s["positions"], dtype=float),
                                 float(s["cutoff"]))


def _check_batch(systems, out, f32=False):
    i, j, D, d, S, off = out
    assert off.shape == (len(systems) + 1,)
    assert off[0] == 0 and off[-1] == len(i)
    for k, s in enumerate(systems):
        a, b = int(off[k]), int(off[k + 1])
        ref = _oracle_system(s)
        assert b - a == len(ref["i"]), f"system {k}: {b - a} pairs, oracle {len(ref['i'])}"
        assert np.array_equal(i[a:b], ref["i"]), f"system {k}: i"
        assert np.array_equal(j[a:b], ref["j"]), f"system {k}: j"
        assert np.array_equal(S[a:b], ref["S"]), f"system {k}: S"
        if f32:
            np.testing.assert_array_equal(D[a:b], ref["D"].astype(np.float32))
            np.testing.assert_array_equal(d[a:b], ref["d"].astype(np.float32))
        else:
            assert np.array_equal(bits(D[a:b]), bits(ref["D"])), f"system {k}: D bits"
            assert np.array_equal(bits(d[a:b]), bits(ref["d"])), f"system {k}: d bits"


def test_uneven_batch_1024_matches_oracle(gpu):
    """The 1024-system uneven workload, every system bit-exact."""
    from matscipy_neighbours_amd import neighbour_list_batch, workloads
    systems = workloads.uneven_batch()
    pos, cells, pbc, rc, nat = _concat(systems)
    print(f"batch: {len(systems)} systems, {len(pos)} atoms", flush=True)
    out = neighbour_list_batch("ijDdS", pos, cells, pbc, rc, nat)
    # neighbour_list_batch returns in the order asked: i, j, D, d, S, offsets
    i, j, D, d, S, off = out
    _check_batch(systems, (i, j, D, d, S, off))
    # boxes narrower than twice the cutoff: one atom is a neighbour through
    # several periodic images (the same (system, i, j) with different S)
    sys_idx = np.repeat(np.arange(len(systems)), np.diff(off))
    key = (sys_idx.astype(np.int64) << 40) | (i.astype(np.int64) << 20) | j.astype(np.int64)
    _, mult = np.unique(key, return_counts=True)
    assert mult.max() >= 2


def test_batch_equals_single_system_builds(gpu):
    """A batch equals the concatenation of single-system builds."""
    from matscipy_neighbours_amd import neighbour_list, neighbour_list_batch, workloads
    systems = workloads.uneven_batch(nsys=40, seed=9)
    pos, cells, pbc, rc, nat = _concat(systems)
    i, j, D, d, S, off = neighbour_list_batch("ijDdS", pos, cells, pbc, rc, nat)
    for k, s in enumerate(systems):
        a, b = int(off[k]), int(off[k + 1])
        r = neighbour_list("ijDdS", positions=s["positions"], cell=s["cell"], pbc=s["pbc"],
                           cutoff=s["cutoff"])
        for got, want in zip((i[a:b], j[a:b], D[a:b], d[a:b], S[a:b]), r):
            assert np.array_equal(bits(got), bits(want)), k


def test_batch_mixed_systems(gpu):
    """Empty systems, mixed periodicity, atoms outside their cell, different
    cutoffs per system, a sparse box and a crowded one."""
    from matscipy_neighbours_amd import neighbour_list_batch
    rng = np.random.default_rng(77)
    systems = []
    for k in range(30):
        n = int(rng.integers(0, 300)) if k % 7 else 0
        edge = rng.uniform(3.0, 25.0)
        cell = edge * (np.eye(3) + rng.uniform(-0.15, 0.15, (3, 3)) * (k % 2))
        pb = rng.integers(0, 2, 3).astype(bool)
        frac = rng.uniform(-0.2, 1.2, (n, 3))
        systems.append(dict(positions=frac @ cell, cell=cell, pbc=pb,
                            cutoff=float(rng.uniform(0.8, 6.0))))
    pos, cells, pbc, rc, nat = _concat(systems)
    i, j, D, d, S, off = neighbour_list_batch("ijDdS", pos, cells, pbc, rc, nat)
    _check_batch(systems, (i, j, D, d, S, off))


def _mixed_small(rng, count, big=0):
    """Systems of 0-300 atoms (and optionally one of `big` atoms), at most
    10 cells per direction, mixed periodicity, atoms up to 20% outside."""
    systems = []
    for k in range(count):
        n = int(rng.integers(0, 300)) if k % 7 else 0
        if big and k == count // 2:
            n = big
        edge = rng.uniform(3.0, 25.0)
        cell = edge * (np.eye(3) + rng.uniform(-0.1, 0.1, (3, 3)) * (k % 2))
        pb = rng.integers(0, 2, 3).astype(bool)
        frac = rng.uniform(-0.2, 1.2, (n, 3))
        systems.append(dict(positions=frac @ cell, cell=cell, pbc=pb,
                            cutoff=float(rng.uniform(edge / 9.0, 6.0 + edge / 9.0))))
    return systems


@pytest.mark.parametrize("big", [0, 3000])
def test_batch_binning_paths(gpu, big):
    """Batches whose systems all have at most 2048 atoms and 2048 cells are
    binned one block per system; one system of 3000 atoms sends the batch
    through the general binning. Both give every system bit-exact pairs,
    with clamped non-periodic directions, atoms outside the cell, empty
    systems, and per-system wave splits (one-cell systems of many atoms)."""
    from matscipy_neighbours_amd import neighbour_list_batch
    rng = np.random.default_rng(31 + big)
    systems = _mixed_small(rng, 40, big)
    for s in systems:
        c = np.asarray(s["cell"])
        faces = abs(np.linalg.det(c)) / np.linalg.norm(np.cross(c[[1, 2, 0]], c[[2, 0, 1]]), axis=1)
        assert np.prod(np.maximum(1, np.floor(faces / s["cutoff"]))) <= 2048
    pos, cells, pbc, rc, nat = _concat(systems)
    print(f"batch: {len(systems)} systems, {len(pos)} atoms, largest {nat.max()}", flush=True)
    out = neighbour_list_batch("ijDdS", pos, cells, pbc, rc, nat)
    _check_batch(systems, out)


def test_batch_f32_output(gpu):
    import numpy as _np
    from matscipy_neighbours_amd import neighbour_list_batch, workloads
    systems = workloads.uneven_batch(nsys=24, seed=13)
    pos, cells, pbc, rc, nat = _concat(systems)
    out = neighbour_list_batch("ijDdS", pos, cells, pbc, rc, nat, out_dtype=_np.float32)
    assert out[2].dtype == np.float32 and out[3].dtype == np.float32
    _check_batch(systems, out, f32=True)


def test_batch_graph_replays(gpu):
    """The captured batch graph, replayed three times with eager work and
    synchronisations in between, moving the atoms between replays: every
    replay equals an eager batch build of the same positions."""
    import torch

    from matscipy_neighbours_amd import _lib, neighbour_list_batch, workloads
    from matscipy_neighbours_amd.neighbours import _batch_geometry
    lib = _lib.load()
    systems = workloads.uneven_batch(nsys=128, seed=21)
    pos, cells, pbc, rc, nat = _concat(systems)
    nsys, offs, co, ce, inv, pb, rcs = _batch_geometry(cells, pbc, rc, nat, None)
    dev = torch.device("cuda:0")
    plan = _lib.Plan(0)
    stream = torch.cuda.current_stream(dev)
    sp = ctypes.c_void_p(stream.cuda_stream)
    P_ = _lib.ptr
    pos_t = torch.from_numpy(pos).to(dev)
    _lib.check(lib.mnl_plan_set_batch(plan.handle, sp, nsys, P_(offs), P_(co), P_(ce), P_(inv),
                                      P_(pb), P_(rcs)))
    n0 = ctypes.c_int64(0)
    _lib.check(lib.mnl_plan_batch_count(plan.handle, sp, P_(pos_t), ctypes.byref(n0)))
    cap = int(n0.value) + 4096
    bufs = dict(i=torch.empty(cap, dtype=torch.int32, device=dev),
                j=torch.empty(cap, dtype=torch.int32, device=dev),
                D=torch.empty((cap, 3), dtype=torch.float64, device=dev),
                d=torch.empty(cap, dtype=torch.float64, device=dev),
                S=torch.empty((cap, 3), dtype=torch.int32, device=dev))
    sys_off = torch.empty(nsys + 1, dtype=torch.int64, device=dev)
    np_dev = torch.zeros(1, dtype=torch.int64, device=dev)
    g = ctypes.c_void_p()
    _lib.check(lib.mnl_plan_batch_graph_create(plan.handle, P_(pos_t), cap, P_(bufs["i"]),
                                               P_(bufs["j"]), P_(bufs["D"]), P_(bufs["d"]),
                                               P_(bufs["S"]), 0, P_(sys_off), P_(np_dev),
                                               ctypes.byref(g)))
    try:
        rng = np.random.default_rng(5)
        for rep in range(3):
            # move the atoms a little (in place: the graph reads this buffer)
            new = pos + rng.normal(0.0, 0.05, pos.shape)
            pos_t.copy_(torch.from_numpy(new))
            scratch = torch.randn(1 << 20, device=dev).sum()  # eager work in between
            torch.cuda.synchronize()
            _lib.check(lib.mnl_graph_launch(g, sp))
            torch.cuda.synchronize()
            P = int(np_dev.item())
            assert P <= cap
            got = [bufs[q][:P].cpu().numpy() for q in "ijDdS"] + [sys_off.cpu().numpy()]
            want = neighbour_list_batch("ijDdS", new, cells, pbc, rc, nat)
            for a, b in zip(got, want):
                assert np.array_equal(bits(a), bits(b)), f"replay {rep}"
            del scratch
    finally:
        lib.mnl_graph_destroy(g)


def test_batch_errors(gpu):
    from matscipy_neighbours_amd import neighbour_list_batch
    # singular cell: inv(cell.T) fails on the host, as in the single-system path
    with pytest.raises(np.linalg.LinAlgError):
        neighbour_list_batch("ij", np.zeros((2, 3)), np.stack([np.eye(3), np.zeros((3, 3))]),
                             True, 1.0, [1, 1])
    # invertible but below the engine's volume threshold (1e-12)
    with pytest.raises(RuntimeError, match="Zero cell volume"):
        neighbour_list_batch("ij", np.zeros((2, 3)), np.stack([np.eye(3), 1e-5 * np.eye(3)]),
                             True, 1.0, [1, 1])
    with pytest.raises(ValueError):
        neighbour_list_batch("ij", np.zeros((3, 3)), np.stack([np.eye(3)] * 2), True, 1.0,
                             [1, 1])
