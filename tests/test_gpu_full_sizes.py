"""The engine at the full workload sizes against the oracle, bit for bit:

* fcc a = 3.615 A, 64^3 cells (1,048,576 atoms), jitter 0.1 A, rc 5 A
  (46.6 M pairs);
* the 200,000-atom two-species triclinic slab, pbc (T,T,F), per-type
  cutoffs 3.0/3.5/4.0 A, through both grid backends (dense and hashed);
* the 131,072-atom sheared bcc lattice at rc 6 A.

Plus the size-independent properties every full list has: D == r[j] - r[i] +
S @ cell, d == |D|, the pair multiset symmetric under (i, j, S) -> (j, i, -S),
i sorted."""

import numpy as np
import pytest

import oracle
from conftest import assert_same_pairs

pytestmark = pytest.mark.gpu


def _engine(w, backend=-1):
    import torch

    from matscipy_neighbours_amd import neighbours as nb
    numbers = w.get("numbers")
    nat = len(w["positions"])
    rc, nums = nb._resolve_cutoff(w["cutoff"], numbers if numbers is not None
                                  else np.ones(nat, np.int32))
    cell = np.ascontiguousarray(w["cell"], dtype=float)
    inv = np.ascontiguousarray(np.linalg.inv(cell.T))
    b = nb._Build(torch.from_numpy(np.ascontiguousarray(w["positions"])).to("cuda:0"),
                  np.zeros(3), cell, inv, np.asarray(w["pbc"], bool), rc, nums, backend=backend)
    out = b.run(31)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}, cell, inv


def _oracle(w, cell, inv):
    return oracle.neighbour_list("ijdDS", np.zeros(3), cell, inv, np.asarray(w["pbc"], bool),
                                 np.ascontiguousarray(w["positions"]), w["cutoff"],
                                 types=w.get("numbers"))


def _properties(w, res, cell):
    i, j, D, d, S = res["i"], res["j"], res["D"], res["d"], res["S"]
    r = w["positions"]
    assert np.all(np.diff(i) >= 0)
    ref_D = r[j] - r[i] + S @ cell
    assert np.max(np.abs(D - ref_D)) < 1e-9
    assert np.max(np.abs(d - np.linalg.norm(D, axis=1))) < 1e-12
    # symmetry: every (i, j, S) has its (j, i, -S)
    key = lambda a, b, s: ((a.astype(np.int64) * len(r) + b) * 64 + (s[:, 0] + 2)) * 64 * 64 + \
        (s[:, 1] + 2) * 64 + (s[:, 2] + 2)
    k1 = np.sort(key(i, j, S))
    k2 = np.sort(key(j, i, -S))
    assert np.array_equal(k1, k2)


@pytest.mark.parametrize("name", ["fcc_1M", "bcc_131k"])
def test_full_size_bitwise(gpu, name):
    from matscipy_neighbours_amd import workloads
    w = workloads.fcc_bulk() if name == "fcc_1M" else workloads.bcc_sheared()
    print(f"{name}: {len(w['positions'])} atoms", flush=True)
    res, cell, inv = _engine(w)
    ref = _oracle(w, cell, inv)
    assert_same_pairs(res, ref, msg=name)
    _properties(w, res, cell)


@pytest.mark.parametrize("backend", [0, 1])
def test_slab_per_type_bitwise(gpu, backend):
    """The sparse slab, per-type cutoffs, dense and hashed grids."""
    from matscipy_neighbours_amd import workloads
    w = workloads.triclinic_slab()
    print(f"slab: {len(w['positions'])} atoms, backend {backend}", flush=True)
    res, cell, inv = _engine(w, backend=backend)
    ref = _oracle(w, cell, inv)
    assert_same_pairs(res, ref, msg=f"slab backend {backend}")
    _properties(w, res, cell)


def test_many_atoms_row_offsets(gpu):
    """4.6 M atoms (more than the 4096 x 1024 counts one scan pass adds up
    in its apply blocks, so the three-kernel scan runs), uniform at 0.01
    atoms/A^3, rc 3 A: the rows follow the per-atom counts (i sorted, each
    atom's row length equal to its count, the total equal to the sum), and
    D, d are consistent with the positions."""
    import torch

    from matscipy_neighbours_amd import neighbours as nb
    n = 4_600_000
    L = (n / 0.01) ** (1.0 / 3.0)
    rng = np.random.default_rng(21)
    r = rng.uniform(0.0, L, (n, 3))
    cell = L * np.eye(3)
    inv = np.ascontiguousarray(np.linalg.inv(cell.T))
    print(f"{n} atoms, box {L:.1f} A", flush=True)
    b = nb._Build(torch.from_numpy(r).to("cuda:0"), np.zeros(3), cell, inv,
                  np.ones(3, bool), 3.0, None)
    out = b.run(nb._quantity_flags("ijdDS"), want_counts=True)
    torch.cuda.synchronize()
    i, j, c = out["i"].cpu().numpy(), out["j"].cpu().numpy(), out["c"].cpu().numpy()
    assert len(i) == int(c.astype(np.int64).sum()) > 0
    assert np.all(np.diff(i) >= 0)
    assert np.array_equal(np.bincount(i, minlength=n), c)
    sel = rng.integers(0, len(i), 200_000)
    D = out["D"].cpu().numpy()[sel]
    S = out["S"].cpu().numpy()[sel]
    ref_D = r[j[sel]] - r[i[sel]] + S @ cell
    assert np.max(np.abs(D - ref_D)) < 1e-9
    d = out["d"].cpu().numpy()[sel]
    assert np.max(np.abs(d - np.linalg.norm(D, axis=1))) < 1e-12
    assert np.all(d < 3.0)
