"""Rows longer than 2^25 pairs: one atom in a periodic unit cube with a
cutoff of 201.5 (the reach is 202 cells each way, 405^3 neighbour-cell
offsets), so that the atom's own row holds 34,262,456 images of itself.

The fill keeps each pair's rank in its row in a word of its own, so a row
may be as long as the int32 offsets allow (src/libneighbours/neighbour_list.cc:69-94
has no per-atom limit below that). The whole list is checked against a
brute-force enumeration of the lattice, in the reference's traversal order
(z outer, y, x inner, each from -202 to 202; neighbour_visit.hh:63-84): with
the unit cell and one atom, S is the lattice vector, D == S exactly and
d == sqrt((Sx^2 + Sy^2) + Sz^2). The dense matrix build of the same input
must report the overflow and write nothing outside its row of K slots.
"""

import ctypes

import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu

RC = 201.5
REACH = 202


def _lattice_images():
    """Every lattice vector S != 0 with (Sx^2 + Sy^2) + Sz^2 < RC^2, in
    traversal order (z slowest, x fastest)."""
    r = np.arange(-REACH, REACH + 1, dtype=np.int64)
    y, x = np.meshgrid(r, r, indexing="ij")
    xy = x * x + y * y
    parts = []
    for z in r:
        keep = (xy + z * z) < RC * RC
        if z == 0:
            keep &= ~((x == 0) & (y == 0))
        parts.append(np.stack([x[keep], y[keep], np.full(int(keep.sum()), z)], axis=1))
    return np.concatenate(parts).astype(np.int32)


def test_row_beyond_2_25_pairs(gpu):
    import torch

    from matscipy_neighbours_amd import neighbour_list

    want_S = _lattice_images()
    npairs = len(want_S)
    assert npairs > 2 ** 25

    pos = torch.tensor([[0.5, 0.5, 0.5]], dtype=torch.float64, device="cuda:0")
    print(f"long row: 1 atom, unit cube, rc {RC}: {npairs} pairs expected", flush=True)
    i, j, d, D, S = neighbour_list("ijdDS", positions=pos, cell=np.eye(3),
                                   pbc=[True, True, True], cutoff=RC)
    torch.cuda.synchronize()
    assert i.shape == (npairs,)
    assert int(i.max()) == 0 and int(i.min()) == 0
    assert int(j.max()) == 0 and int(j.min()) == 0
    S = S.cpu().numpy()
    assert np.array_equal(S, want_S)
    want_D = want_S.astype(np.float64)
    assert np.array_equal(bits(D.cpu().numpy()), bits(want_D))
    r2 = (want_D[:, 0] * want_D[:, 0] + want_D[:, 1] * want_D[:, 1]) + want_D[:, 2] * want_D[:, 2]
    assert np.array_equal(bits(d.cpu().numpy()), bits(np.sqrt(r2)))


def test_matrix_of_long_row_stays_in_its_row(gpu):
    import torch

    from matscipy_neighbours_amd import _lib

    lib = _lib.load()
    K, guard = 4096, 4096
    sentinel = -12345
    idx = torch.full((K + guard,), sentinel, dtype=torch.int32, device="cuda:0")
    dist = torch.full((K + guard, 3), float(sentinel), dtype=torch.float64, device="cuda:0")
    cnt = torch.full((2,), sentinel, dtype=torch.int32, device="cuda:0")
    pos = torch.tensor([[0.5, 0.5, 0.5]], dtype=torch.float64, device="cuda:0")
    co = np.zeros(3)
    ce = np.eye(3)
    inv = np.eye(3)
    pbc = np.ones(3, np.int32)
    plan = _lib.Plan(0)
    ovf = ctypes.c_int32(-1)
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(lib.mnl_plan_matrix(plan.handle, sp, _lib.ptr(co), _lib.ptr(ce), _lib.ptr(inv),
                                   _lib.ptr(pbc), 1, _lib.ptr(pos), RC, None, None, 0, None,
                                   -1, K, _lib.ptr(idx), _lib.ptr(dist), _lib.ptr(cnt),
                                   ctypes.byref(ovf)))
    torch.cuda.synchronize()
    want_S = _lattice_images()
    assert ovf.value == 1
    idx = idx.cpu().numpy()
    dist = dist.cpu().numpy()
    cnt = cnt.cpu().numpy()
    assert cnt[0] == len(want_S) and cnt[1] == sentinel
    assert np.all(idx[:K] == 0)
    assert np.array_equal(bits(dist[:K]), bits(want_S[:K].astype(np.float64)))
    assert np.all(idx[K:] == sentinel)
    assert np.all(dist[K:] == float(sentinel))
