"""The engine at the full workload sizes against the oracle, bit for bit:

* fcc a = 3.615 A, 64^3 cells (1,048,576 atoms), jitter 0.1 A, rc 5 A
  (46.6 M pairs);
* the 200,000-atom two-species triclinic slab, pbc (T,T,F), per-type
  cutoffs 3.0/3.5/4.0 A, through both grid backends (dense and hashed);
* the 131,072-atom sheared bcc lattice at rc 6 A.

Each full list is also pinned to the compiled reference itself: the SHA-256
digests of its i, j, d, D and S written by tests/golden/make_golden.py from
the reference CPU core on the same seeded inputs (whose positions digest is
checked first, so a generator change is reported as such).

Plus the size-independent properties every full list has: D == r[j] - r[i] +
S @ cell, d == |D|, the pair multiset symmetric under (i, j, S) -> (j, i, -S),
i sorted."""

import numpy as np
import pytest

import oracle
from conftest import assert_same_pairs
from test_oracle import INDEX, digest

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


def _reference_digests(name, w, res):
    """Bit-exactness against the compiled reference's full-size outputs."""
    entry = INDEX["full_sizes"][name]
    assert digest(np.ascontiguousarray(w["positions"], dtype=np.float64)) == \
        entry["positions_digest"], f"{name}: the seeded generator no longer gives the same input"
    assert len(res["i"]) == entry["npairs"], name
    for q in "ijdDS":
        assert digest(res[q]) == entry["digest"][q], (name, q)


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
    _reference_digests(name, w, res)
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
    _reference_digests("slab_200k", w, res)
    ref = _oracle(w, cell, inv)
    assert_same_pairs(res, ref, msg=f"slab backend {backend}")
    _properties(w, res, cell)


def test_many_atoms_row_offsets(gpu):
    """4.6 M atoms (more than the 2048 x 1024 counts one scan pass adds up
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


def test_beyond_int32_pairs(gpu):
    """More pairs than an int32 can index (the reference's index_t limits
    its lists to 2^31 - 1, src/libneighbours/types.hh:28; this engine's
    counts, row offsets and output positions are int64): 26 M atoms of a
    uniform gas at 0.1 atoms/A^3, rc 6 A, about 2.35 G pairs and 122 GB of
    float64 ijdDS output. Checked on the device:

    * the total exceeds 2^31 and equals the sum of the per-atom counts;
    * i is sorted and every row starts at the prefix sum of the counts before
      it and ends where the next starts (so every row holds its own atom's
      pairs, including the rows written past position 2^31);
    * j is a valid atom, S is in {-1, 0, 1} (rc is far below half the box);
    * on 1 M random pairs plus the pairs around positions 2^31 and at the
      end, D == r[j] - r[i] + S L and d == |D| < rc;
    * 64 rows (48 of them past position 2^31, 16 of those of atoms near a
      face, whose pairs cross the boundary) hold exactly the (j, S) that a
      brute-force minimum-image search finds within rc, up to pairs within
      1e-9 A of rc, whose membership the two roundings may decide apart.

    first_neighbours, whose seeds are int32 as the reference's are, refuses
    the list."""
    import torch

    from matscipy_neighbours_amd import neighbours as nb
    free, total = torch.cuda.mem_get_info()
    if free < 160 * 2**30:
        pytest.skip(f"needs 160 GiB of free device memory, {free / 2**30:.0f} GiB free")
    n = 26_000_000
    rc = 6.0
    L = (n / 0.1) ** (1.0 / 3.0)
    rng = np.random.default_rng(2024)
    r_host = rng.uniform(0.0, L, (n, 3))
    cell = L * np.eye(3)
    inv = np.ascontiguousarray(np.linalg.inv(cell.T))
    print(f"{n} atoms, box {L:.1f} A, rc {rc}", flush=True)
    r = torch.from_numpy(r_host).to("cuda:0")
    del r_host
    b = nb._Build(r, np.zeros(3), cell, inv, np.ones(3, bool), rc, None)
    out = b.run(nb._quantity_flags("ijdDS"), want_counts=True)
    torch.cuda.synchronize()
    i, j, D, d, S, c = (out[k] for k in ("i", "j", "D", "d", "S", "c"))
    P = i.numel()
    print(f"{P} pairs", flush=True)
    assert P > 2**31 + 10**8
    assert j.numel() == d.numel() == P and D.shape == S.shape == (P, 3)

    off = torch.zeros(n + 1, dtype=torch.int64, device=r.device)
    off[1:] = torch.cumsum(c.to(torch.int64), 0)
    assert int(off[-1]) == P
    assert bool((i[1:] >= i[:-1]).all())
    full = torch.nonzero(c > 0).squeeze(1)
    assert bool((i[off[full]] == full.to(torch.int32)).all())
    assert bool((i[off[full + 1] - 1] == full.to(torch.int32)).all())
    assert bool((j >= 0).all()) and bool((j < n).all())
    assert bool((S >= -1).all()) and bool((S <= 1).all())
    print("rows and ranges checked", flush=True)

    g = torch.Generator(device=r.device).manual_seed(5)
    sel = torch.cat([
        torch.randint(0, P, (1_000_000,), generator=g, device=r.device),
        torch.arange(2**31 - 50_000, 2**31 + 50_000, device=r.device),
        torch.arange(P - 50_000, P, device=r.device)])
    Ds = D[sel]
    ref_D = r[j[sel].long()] - r[i[sel].long()] + S[sel].double() * L
    assert float((Ds - ref_D).abs().max()) < 1e-9
    ds = d[sel]
    assert float((ds - torch.linalg.norm(Ds, dim=1)).abs().max()) < 1e-12
    assert bool((ds < rc).all())
    del Ds, ref_D, ds
    print("sampled pairs checked", flush=True)

    # rows past position 2^31, a third of them of atoms within rc of a face
    # (the only ones with nonzero shifts), and rows anywhere
    late = full[off[full] > 2**31]
    edge = torch.minimum(r, L - r).min(dim=1).values < rc
    late_edge = late[edge[late]]
    pick = lambda src, k: src[torch.randint(0, src.numel(), (k,), generator=g,
                                            device=r.device)]
    rows = torch.cat([pick(late, 32), pick(late_edge, 16), pick(full, 16)])
    for a in rows.tolist():
        dr = r - r[a]
        s = -torch.round(dr / L)
        dist = torch.linalg.norm(dr + s * L, dim=1)
        dist[a] = float("inf")
        key = lambda jj, ss: (jj.long() * 4 + (ss[:, 0].long() + 1)) * 16 + \
            (ss[:, 1].long() + 1) * 4 + (ss[:, 2].long() + 1)
        lo, hi = int(off[a]), int(off[a + 1])
        got = set(key(j[lo:hi], S[lo:hi]).tolist())
        inner = torch.nonzero(dist < rc - 1e-9).squeeze(1)
        outer = torch.nonzero(dist < rc + 1e-9).squeeze(1)
        want_in = set(key(inner, s[inner]).tolist())
        want_out = set(key(outer, s[outer]).tolist())
        assert want_in <= got <= want_out, f"row {a}"
    print("brute-force rows checked", flush=True)

    with pytest.raises(RuntimeError, match="2\\^31"):
        nb.first_neighbours(n, i)


def test_full_size_host_arrays(gpu):
    """fcc 1M through the public API with numpy in and out: the 2.4 GB result
    comes back in one transfer (pageable above 256 MiB), bit for bit against
    the oracle."""
    from matscipy_neighbours_amd import neighbour_list, workloads
    w = workloads.fcc_bulk()
    got = neighbour_list("ijdDS", positions=w["positions"], cell=w["cell"], pbc=w["pbc"],
                         cutoff=w["cutoff"])
    assert all(isinstance(a, np.ndarray) for a in got)
    cell = np.ascontiguousarray(w["cell"], dtype=float)
    inv = np.ascontiguousarray(np.linalg.inv(cell.T))
    ref = _oracle(w, cell, inv)
    assert_same_pairs(dict(zip("ijdDS", got)), ref, msg="fcc 1M host arrays")
