"""The autotuner and the launch shapes it chooses between. A shape (home
atoms of a cell dealt to 1..64 waves, fill windows of 4 or 5 chunks) changes
how the work is dealt to waves and never the outputs: every shape is checked
bit for bit against the oracle, on a uniform system (dense and hashed grids),
per-atom and per-type cutoffs and a batch; the tuned shape is then used by
later builds of the same configuration class."""

import ctypes

import numpy as np
import pytest

import oracle
from conftest import assert_same_pairs, bits

pytestmark = pytest.mark.gpu

SHAPES = [(s, c) for s in (1, 2, 4, 8, 64) for c in (4, 5)]


def _plan_build(lib, _lib, plan, sp, torch, pos, cell, pbc, cutoff, numbers, backend):
    from matscipy_neighbours_amd import neighbours as nb
    pos_t = torch.from_numpy(np.ascontiguousarray(pos)).cuda()
    co, ce, inv, pb, nums = nb._host_metadata(cell, pbc, numbers, None, len(pos),
                                              positions=pos)
    rc, nums = nb._resolve_cutoff(cutoff, nums)
    b = nb._Build(pos_t, co, ce, inv, pb, rc, nums, backend=backend)
    n = ctypes.c_int64(0)
    P_ = _lib.ptr
    _lib.check(lib.mnl_plan_count(plan.handle, sp, P_(b.co), P_(b.ce), P_(b.inv), P_(b.pbc),
                                  b.nat, P_(b.pos), b.rc, P_(b.per_atom_t), P_(b.per_type_sq),
                                  b.nc, P_(b.types_t), backend, ctypes.byref(n)))
    P = int(n.value)
    out = dict(i=torch.empty(P, dtype=torch.int32, device="cuda"),
               j=torch.empty(P, dtype=torch.int32, device="cuda"),
               D=torch.empty((P, 3), dtype=torch.float64, device="cuda"),
               d=torch.empty(P, dtype=torch.float64, device="cuda"),
               S=torch.empty((P, 3), dtype=torch.int32, device="cuda"))
    _lib.check(lib.mnl_plan_fill(plan.handle, sp, *(P_(out[q]) for q in "ijDdS")))
    torch.cuda.synchronize()
    return {q: v.cpu().numpy() for q, v in out.items()}


def _oracle(pos, cell, pbc, cutoff, numbers):
    cell = np.asarray(cell, float)
    inv = np.ascontiguousarray(np.linalg.inv(cell.T))
    types = numbers if np.asarray(cutoff).ndim == 2 else None
    return oracle.neighbour_list("ijdDS", np.zeros(3), cell, inv, pbc, np.ascontiguousarray(pos),
                                 cutoff, types=types)


@pytest.mark.parametrize("case", ["uniform_dense", "uniform_hashed", "per_atom", "per_type"])
def test_every_launch_shape_gives_the_same_pairs(gpu, case):
    import torch

    from matscipy_neighbours_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(101)
    n = 2500
    cell = np.array([[13.0, 0.0, 0.0], [1.0, 14.0, 0.0], [0.5, 0.3, 15.0]])
    pos = rng.uniform(0.0, 1.0, (n, 3)) @ cell
    numbers, backend = None, 0
    cutoff = 2.6
    if case == "uniform_hashed":
        backend = 1
    elif case == "per_atom":
        cutoff = rng.uniform(1.0, 1.3, n)
    elif case == "per_type":
        numbers = rng.integers(0, 2, n).astype(np.int32)
        cutoff = np.array([[2.2, 2.6], [2.6, 2.4]])
    ref = _oracle(pos, cell, True, cutoff, numbers)
    plan = _lib.Plan(0)
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for split, chunks in SHAPES:
        _lib.check(lib.mnl_plan_set_launch_shape(plan.handle, split, chunks))
        got = _plan_build(lib, _lib, plan, sp, torch, pos, cell, True, cutoff, numbers, backend)
        assert_same_pairs(got, ref, msg=f"{case} split={split} chunks={chunks}")
    _lib.check(lib.mnl_plan_set_launch_shape(plan.handle, 0, 0))


def test_launch_shape_arguments(gpu):
    from matscipy_neighbours_amd import _lib
    lib = _lib.load()
    plan = _lib.Plan(0)
    assert lib.mnl_plan_set_launch_shape(plan.handle, 65, 4) == -1
    assert lib.mnl_plan_set_launch_shape(plan.handle, 2, 3) == -1
    assert lib.mnl_plan_set_launch_shape(plan.handle, 2, 5) == 0


def test_autotune_single_system(gpu):
    """autotune picks one of the candidate shapes; later builds of the same
    class use it and still match the oracle bit for bit."""
    from matscipy_neighbours_amd import _lib, autotune, neighbour_list
    lib = _lib.load()
    rng = np.random.default_rng(102)
    cell = np.diag([20.0, 21.0, 22.0])
    pos = rng.uniform(0.0, 1.0, (6000, 3)) @ cell
    try:
        r = autotune(positions=pos, cell=cell, pbc=True, cutoff=2.5)
        assert r["split"] in (1, 2, 4, 8) and r["chunks"] in (4, 5) and r["ms"] > 0
        got = neighbour_list("ijdDS", positions=pos, cell=cell, pbc=True, cutoff=2.5)
        ref = _oracle(pos, cell, True, 2.5, None)
        assert_same_pairs(dict(zip("ijdDS", got)), ref, msg=f"tuned {r}")
    finally:
        lib.mnl_autotune_clear()


def test_autotune_batch(gpu):
    from matscipy_neighbours_amd import _lib, autotune_batch, neighbour_list_batch, workloads
    lib = _lib.load()
    systems = workloads.uneven_batch(nsys=128, seed=41)
    pos = np.concatenate([s["positions"] for s in systems])
    cells = np.stack([s["cell"] for s in systems])
    nat = np.array([len(s["positions"]) for s in systems])
    try:
        before = neighbour_list_batch("ijDdS", pos, cells, True, 5.0, nat)
        r = autotune_batch(pos, cells, True, 5.0, nat)
        assert r["split"] in (1, 2, 4, 8) and r["chunks"] in (4, 5)
        after = neighbour_list_batch("ijDdS", pos, cells, True, 5.0, nat)
        for a, b in zip(before, after):
            assert np.array_equal(bits(a), bits(b)), f"tuned {r}"
    finally:
        lib.mnl_autotune_clear()
