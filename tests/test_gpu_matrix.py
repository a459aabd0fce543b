"""Dense neighbour matrix on the GPU (one traversal, no pair list) against
the compiled reference's neighbour_matrix fixtures and the oracle's
restatement (the pair list reshaped into rows of K slots), bit for bit;
the overflow flag and the Python error for too small a capacity."""

import ctypes

import numpy as np
import pytest

import oracle
from conftest import bits
from test_oracle import ARR, INDEX, case_inputs

pytestmark = pytest.mark.gpu


def _engine_matrix(pos, cell, inv, pbc, cutoff, types, K, backend=-1):
    import torch

    from matscipy_neighbours_amd import _lib
    from matscipy_neighbours_amd import neighbours as nb
    lib = _lib.load()
    n = len(pos)
    rc, nums = nb._resolve_cutoff(cutoff, types if types is not None else np.ones(n, np.int32))
    b = nb._Build(torch.from_numpy(np.ascontiguousarray(pos)).to("cuda:0"), np.zeros(3),
                  np.ascontiguousarray(cell), np.ascontiguousarray(inv), pbc, rc, nums)
    plan = _lib.Plan(0)
    idx = torch.full((n, K), -7, dtype=torch.int32, device="cuda:0")
    dist = torch.full((n, K, 3), -7.0, dtype=torch.float64, device="cuda:0")
    cnt = torch.full((n,), -7, dtype=torch.int32, device="cuda:0")
    ovf = ctypes.c_int32(-1)
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(lib.mnl_plan_matrix(plan.handle, sp, _lib.ptr(b.co), _lib.ptr(b.ce),
                                   _lib.ptr(b.inv), _lib.ptr(b.pbc), n, _lib.ptr(b.pos), b.rc,
                                   _lib.ptr(b.per_atom_t), _lib.ptr(b.per_type_sq), b.nc,
                                   _lib.ptr(b.types_t), backend, K, _lib.ptr(idx),
                                   _lib.ptr(dist), _lib.ptr(cnt), ctypes.byref(ovf)))
    torch.cuda.synchronize()
    return idx.cpu().numpy(), dist.cpu().numpy(), cnt.cpu().numpy(), bool(ovf.value)


def test_matrix_matches_golden(gpu):
    for c in INDEX["post"]["neighbour_matrix"]:
        name, K = c["case"], c["K"]
        pos, cell, inv, pbc, cutoff, types = case_inputs(name)
        for backend in (0, 1):
            idx, dist, cnt, ovf = _engine_matrix(pos, cell, inv, pbc, cutoff, types, K, backend)
            key = f"matrix/{name}/K{K}"
            assert np.array_equal(idx, ARR[f"{key}/idx"]), (key, backend)
            assert np.array_equal(bits(dist), bits(ARR[f"{key}/dist"])), (key, backend)
            assert np.array_equal(cnt, ARR[f"{key}/count"]), (key, backend)
            assert ovf == c["overflow"], (key, backend)


@pytest.mark.parametrize("seed", range(4))
def test_matrix_matches_oracle(gpu, seed):
    rng = np.random.default_rng(300 + seed)
    n = int(rng.integers(50, 3000))
    cell = np.diag(rng.uniform(6, 20, 3)) + rng.uniform(-1, 1, (3, 3)) * (seed % 2)
    pbc = rng.integers(0, 2, 3).astype(bool)
    pos = rng.uniform(-0.1, 1.1, (n, 3)) @ cell
    inv = np.ascontiguousarray(np.linalg.inv(cell.T))
    rc = float(rng.uniform(1.5, 4.0))
    ref = oracle.neighbour_matrix(np.zeros(3), cell, inv, pbc, pos, rc, 1)
    # K below the largest row (overflow), at it, and above it by a few and
    # by more than 64 slots (rows whose zeroed tail takes several passes)
    for K in (max(1, int(ref[2].max()) - 3), int(ref[2].max()), int(ref[2].max()) + 5,
              int(ref[2].max()) + 70):
        want = oracle.neighbour_matrix(np.zeros(3), cell, inv, pbc, pos, rc, K)
        got = _engine_matrix(pos, cell, inv, pbc, rc, None, K)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[2], want[2])
        assert np.array_equal(bits(got[1]), bits(want[1]))
        assert got[3] == want[3]


def test_python_neighbour_matrix(gpu):
    from matscipy_neighbours_amd import neighbour_list, neighbour_matrix
    rng = np.random.default_rng(8)
    pos = rng.uniform(0, 10, (500, 3))
    idx, dist, cnt = neighbour_matrix(positions=pos, cell=10 * np.eye(3), pbc=True,
                                      cutoff=2.5, max_neighbours=64)
    assert idx.shape == (500, 64) and dist.shape == (500, 64, 3) and cnt.shape == (500,)
    i, j, D = neighbour_list("ijD", positions=pos, cell=10 * np.eye(3), pbc=True, cutoff=2.5)
    assert np.array_equal(cnt, np.bincount(i, minlength=500))
    mask = np.arange(64)[None, :] < cnt[:, None]
    assert np.array_equal(idx[mask], j) and np.array_equal(dist[mask], D)
    assert not idx[~mask].any() and not dist[~mask].any()
    with pytest.raises(ValueError, match="too small"):
        neighbour_matrix(positions=pos, cell=10 * np.eye(3), pbc=True, cutoff=2.5,
                         max_neighbours=int(cnt.max()) - 1)
