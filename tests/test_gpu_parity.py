"""HIP engine vs the oracle: bit-exact pairs, order, D, d and S.

Every case runs the engine through the C ABI (libmnl.so via the Python
binding) on the GPU and the oracle (the CPU restatement of the reference
path) on the same inputs, with the same inv_cell bytes, and requires
identical arrays bit for bit. Cases follow the reference's own tests
(tests/test_extension.py, tests/test_core.cc, tests/test_neighbour_list_gpu.cc)
plus boundary stress (atoms on cell faces, pairs exactly at the cutoff),
tiny cells (multi-image shifts), both grid backends and random triclinic
fuzz.
"""

import numpy as np
import pytest

import oracle
from conftest import assert_same_pairs

pytestmark = pytest.mark.gpu


def _engine(quantities, positions, cell, pbc, cutoff, numbers=None, backend=-1,
            f32=False):
    import torch
    from matscipy_neighbours_amd import neighbours as nb
    pos = np.ascontiguousarray(positions, dtype=float)
    co, ce, inv, pb, nums = nb._host_metadata(cell, pbc, numbers, None, len(pos),
                                              positions=pos)
    rc, nums = nb._resolve_cutoff(cutoff, nums)
    b = nb._Build(torch.from_numpy(pos).to("cuda:0"), co, ce, inv, pb, rc, nums,
                  backend=backend)
    out = b.run(nb._quantity_flags(quantities), f32=f32, want_counts=True)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


def _oracle(quantities, positions, cell, pbc, cutoff, numbers=None):
    pos = np.ascontiguousarray(positions, dtype=float)
    cell = np.asarray(cell, dtype=float)
    inv = np.ascontiguousarray(np.linalg.inv(cell.T))
    types = None
    if numbers is not None and np.asarray(cutoff).ndim == 2:
        types = numbers
    return oracle.neighbour_list(quantities, np.zeros(3), cell, inv, pbc, pos, cutoff,
                                 types=types, with_counts=True)


def _check(positions, cell, pbc, cutoff, numbers=None, backends=(-1,), msg=""):
    ref = _oracle("ijdDS", positions, cell, pbc, cutoff, numbers)
    for be in backends:
        got = _engine("ijdDS", positions, cell, pbc, cutoff, numbers, backend=be)
        assert_same_pairs(got, ref, msg=f"{msg} backend={be}")
        assert np.array_equal(got["c"], ref["counts"]), f"{msg} counts"
    return ref


# ------------------------------------------------------------ known answers

def test_unit_cube_known_answers(gpu):
    """tests/test_core.cc:32-62, tests/test_extension.py:102-119."""
    pos = np.array([[0.5, 0.5, 0.5]])
    ref = _check(pos, np.eye(3), True, 1.1)
    assert len(ref["i"]) == 6
    assert np.array_equal(ref["D"], ref["S"].astype(float))
    assert len(_check(pos, np.eye(3), True, 1.5)["i"]) == 18
    assert len(_check(pos, np.eye(3), False, 1.1)["i"]) == 0
    assert len(_check(pos, np.eye(3), [True, False, False], 1.1)["i"]) == 2
    assert len(_check(pos, np.eye(3), [True, False, True], 1.1)["i"]) == 4


def test_simple_cubic_2x2x2(gpu):
    """tests/test_core.cc:75-122: 8 atoms, 48 pairs."""
    g = np.array([[x, y, z] for x in range(2) for y in range(2) for z in range(2)], float)
    ref = _check(g + 0.5, 2 * np.eye(3), True, 1.1)
    assert len(ref["i"]) == 48


def test_per_atom_and_per_type(gpu):
    """tests/test_extension.py:211-231."""
    cell = 10.0 * np.eye(3)
    pos = np.array([[0.0, 0, 0], [1.0, 0, 0], [2.5, 0, 0]])
    ref = _check(pos, cell, False, np.array([0.4, 0.4, 1.2]))
    assert sorted(zip(ref["i"].tolist(), ref["j"].tolist())) == [(1, 2), (2, 1)]
    pos = np.array([[0.0, 0, 0], [1.0, 0, 0], [2.0, 0, 0]])
    types = np.array([0, 1, 0], dtype=np.int32)
    ref = _check(pos, cell, False, np.array([[0.5, 1.5], [1.5, 0.5]]), numbers=types)
    assert sorted(zip(ref["i"].tolist(), ref["j"].tolist())) == [(0, 1), (1, 0), (1, 2),
                                                                  (2, 1)]


def test_per_type_out_of_range_types_keep_global_test(gpu):
    """Types outside the matrix fall back to the global cutoff
    (neighbour_visit.hh:100-108)."""
    rng = np.random.default_rng(11)
    pos = rng.uniform(0, 8, (300, 3))
    types = rng.integers(-1, 4, 300).astype(np.int32)
    _check(pos, 8 * np.eye(3), True, np.array([[1.0, 1.5, 1.2], [1.5, 2.0, 0.9],
                                               [1.2, 0.9, 1.7]]), numbers=types)


@pytest.mark.parametrize("nt", [11, 12])
def test_per_type_many_types(gpu, nt):
    """11 types (121 table entries, the largest table kept in LDS, where home
    atoms are taken class by class) and 12 (the table read from global
    memory), with types -1 and nt outside the table, symmetric tables with
    entries down to 0.3 rc; cell by cell the home atoms of one cell carry
    many classes."""
    rng = np.random.default_rng(100 + nt)
    n = 3000
    pos = rng.uniform(0, 14.0, (n, 3))
    types = rng.integers(-1, nt + 1, n).astype(np.int32)
    t = rng.uniform(0.6, 2.0, (nt, nt))
    table = (t + t.T) / 2
    _check(pos, 14.0 * np.eye(3), True, table, numbers=types, backends=(0, 1),
           msg=f"{nt} types")


@pytest.mark.parametrize("L", [8.0, 2.5])
def test_per_type_asymmetric_table(gpu, L):
    """A per-type matrix with cut[a][b] != cut[b][a]: the pair (i, j) uses
    the entry [type i][type j] (neighbour_visit.hh:100-108), so (i, j) and
    (j, i) can differ and the count must visit every offset. Also in a box
    smaller than the cutoff (multi-image shifts)."""
    rng = np.random.default_rng(5)
    n = 300 if L > 4 else 12
    pos = rng.uniform(0, L, (n, 3))
    types = rng.integers(-1, 3, n).astype(np.int32)
    cut = np.array([[1.0, 1.6, 0.7], [0.9, 2.0, 1.4], [1.8, 0.6, 1.7]])
    ref = _check(pos, L * np.eye(3), True, cut, numbers=types, backends=(0, 1))
    i, j = ref["i"], ref["j"]
    pairs = set(zip(i.tolist(), j.tolist()))
    assert any((b, a) not in pairs for a, b in pairs), "table asymmetry not exercised"


# ------------------------------------------------------------ reference scenarios

@pytest.mark.parametrize("pbc", [True, False, (True, False, True), (False, True, False)])
def test_cubic_random(gpu, pbc):
    """tests/test_extension.py:126-134."""
    rng = np.random.default_rng(1234)
    _check(rng.uniform(0, 5.0, (40, 3)), 5.0 * np.eye(3), pbc, 1.7, backends=(0, 1))


@pytest.mark.parametrize("pbc", [True, False, (True, True, False)])
def test_triclinic_atoms_outside(gpu, pbc):
    """tests/test_extension.py:137-148."""
    rng = np.random.default_rng(7)
    cell = np.array([[4.0, 0.0, 0.0], [1.3, 3.7, 0.0], [0.8, 0.6, 4.2]])
    pos = rng.uniform(-0.3, 1.3, (30, 3)) @ cell
    _check(pos, cell, pbc, 1.6, backends=(0, 1))


def test_atoms_far_outside(gpu):
    """tests/test_extension.py:163-174, including the non-periodic shift quirk
    for atoms more than a box length outside."""
    rng = np.random.default_rng(99)
    cell = 6.0 * np.eye(3)
    pos = rng.uniform(0, 6.0, (25, 3))
    pos[3] += cell[0]
    pos[7] -= 2 * cell[1]
    pos[11] += 3 * cell[2]
    for pbc in (True, False, (True, False, True)):
        _check(pos, cell, pbc, 1.5, backends=(0, 1), msg=str(pbc))


@pytest.mark.parametrize("pbc", [True, False, (True, False, True)])
def test_vacuum_box(gpu, pbc):
    """tests/test_extension.py:151-160 (dense by the reference rule; also
    forced through the hashed grid)."""
    rng = np.random.default_rng(3)
    _check(rng.uniform(0, 6.0, (250, 3)), 120.0 * np.eye(3), pbc, 1.5, backends=(-1, 0, 1))


@pytest.mark.parametrize("n,L,rc,seed", [(2000, 12.0, 1.0, 1), (500, 3.5, 1.0, 2),
                                         (3000, 15.0, 2.0, 3)])
def test_gpu_suite_scenarios(gpu, n, L, rc, seed):
    """tests/test_neighbour_list_gpu.cc:95-110 (there compared as multisets;
    here bit for bit, order included)."""
    rng = np.random.default_rng(seed)
    _check(rng.uniform(0, L, (n, 3)), L * np.eye(3), True, rc, backends=(0, 1))


def test_sparse_vacuum_hashed(gpu):
    """tests/test_neighbour_list_gpu.cc:120-127: 4000 atoms in a box 40x
    larger than the cluster: the reference rule selects the hashed grid."""
    rng = np.random.default_rng(11)
    Lc = (4000 / 12.0) ** (1 / 3)
    ref = _oracle("ijdDS", rng.uniform(0, Lc, (4000, 3)), 40 * Lc * np.eye(3), True, 1.0)
    assert ref["_sparse"]
    rng = np.random.default_rng(11)
    _check(rng.uniform(0, Lc, (4000, 3)), 40 * Lc * np.eye(3), True, 1.0, backends=(-1, 0))


def test_cfg1_uniform_1000(gpu):
    """1000 atoms uniform in a 10 A periodic cube, rc 2.5 (~65k pairs)."""
    from matscipy_neighbours_amd import workloads
    w = workloads.uniform_box()
    ref = _check(w["positions"], w["cell"], w["pbc"], w["cutoff"], backends=(0, 1))
    assert 60000 < len(ref["i"]) < 70000


# ------------------------------------------------------------ boundary stress

@pytest.mark.parametrize("rc", [1.0, np.sqrt(2.0), np.sqrt(3.0), 2.0])
def test_lattice_at_shell_distance(gpu, rc):
    """Simple-cubic lattice with rc exactly at a shell distance: every pair at
    r2 == rc^2 must be excluded by the same rounding as the reference."""
    g = np.arange(6, dtype=float)
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    _check(pos, 6 * np.eye(3), True, float(rc), backends=(0, 1))


@pytest.mark.parametrize("pbc", [True, (True, False, True)])
def test_pairs_within_ulps_of_the_cutoff(gpu, pbc):
    """Simple-cubic lattice at spacing a = rc = 1.7 with every coordinate
    nudged by up to 3 ulps: hundreds of nearest-neighbour pairs have r2
    within a few ulps of rc^2, on both sides. The count's fast test cannot
    decide those and re-tests them in the reference's operation order; the
    counts and rows must still be the oracle's."""
    rng = np.random.default_rng(77)
    a = 1.7
    g = np.arange(7, dtype=float) * a
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos = pos + rng.integers(-3, 4, pos.shape) * np.spacing(pos + a)
    cell = 7 * a * np.eye(3)
    # nearest-neighbour pairs along the axes inside the box: within 1e-13 of rc^2
    d = pos[:, None, :] - pos[None, :, :]
    r2 = np.einsum("ijk,ijk->ij", d, d)
    near = np.abs(r2 - a * a) < 1e-13
    assert near.sum() > 1000
    assert (r2[near] < a * a).any() and (r2[near] >= a * a).any()
    ref = _check(pos, cell, pbc, a, backends=(0, 1), msg=str(pbc))
    assert 0 < len(ref["i"])


def test_multi_image_pairs_at_the_cutoff(gpu):
    """Two atoms in a 1.7 A cubic box with rc exactly two and three box
    lengths: images at r2 == rc^2 and within ulps of it, reached through
    cell offsets of several box lengths."""
    cell = 1.7 * np.eye(3)
    pos = np.array([[0.0, 0.0, 0.0], [0.85, 0.85, 0.85 + 2 * np.spacing(0.85)]])
    for rc in (3.4, 5.1):
        _check(pos, cell, True, rc, backends=(0, 1), msg=f"rc={rc}")


def test_atoms_on_cell_faces(gpu):
    """Atoms exactly on bin faces k*bin (floor of the cell index decides)."""
    cell = 8.0 * np.eye(3)
    g = np.arange(0, 8.0, 2.0)
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos = np.concatenate([pos, pos + 1.0, pos - 8.0, pos + 8.0])
    for pbc in (True, False, (False, True, True)):
        _check(pos, cell, pbc, 2.0, backends=(0, 1), msg=str(pbc))


@pytest.mark.parametrize("cut", [0.7, 1.5, 2.5, 3.3])
def test_tiny_cells_multi_image(gpu, cut):
    """Boxes narrower than the cutoff: n = 1, several images per neighbour
    and |S| >= 2 (nx = ceil(rc / width))."""
    rng = np.random.default_rng(21)
    cell = np.array([[1.0, 0.0, 0.0], [0.2, 1.1, 0.0], [0.1, -0.3, 0.9]])
    pos = rng.uniform(0, 1, (5, 3)) @ cell
    for pbc in (True, (True, True, False)):
        ref = _check(pos, cell, pbc, cut, backends=(0, 1), msg=str(pbc))
        if cut > 2.0 and pbc is True:
            assert np.abs(ref["S"]).max() >= 2


def test_empty_and_single(gpu):
    _check(np.zeros((0, 3)), 5 * np.eye(3), True, 1.0)
    _check(np.array([[1.0, 2.0, 3.0]]), 5 * np.eye(3), True, 1.0)


def test_big_cell_occupancy(gpu):
    """More than 64 atoms per cell (home-atom groups of 64)."""
    rng = np.random.default_rng(5)
    pos = rng.uniform(0, 4.0, (900, 3))
    _check(pos, 4.0 * np.eye(3), True, 1.9, backends=(0, 1))


@pytest.mark.parametrize("seed", range(12))
def test_random_triclinic_fuzz(gpu, seed):
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.integers(1, 400))
    cell = np.diag(rng.uniform(2.0, 12.0, 3)) + rng.uniform(-1.5, 1.5, (3, 3)) * (seed % 3 > 0)
    pbc = rng.integers(0, 2, 3).astype(bool)
    pos = rng.uniform(-0.4, 1.4, (n, 3)) @ cell
    cutoff = float(rng.uniform(0.3, 4.0))
    _check(pos, cell, pbc, cutoff, backends=(0, 1), msg=f"seed={seed}")


def test_fp32_output_option(gpu):
    """float32 D and d: same pairs as float64, values within 2e-6 relative."""
    from matscipy_neighbours_amd import workloads
    w = workloads.uniform_box(seed=9)
    ref = _oracle("ijdDS", w["positions"], w["cell"], w["pbc"], w["cutoff"])
    got = _engine("ijdDS", w["positions"], w["cell"], w["pbc"], w["cutoff"], f32=True)
    for q in "ijS":
        assert np.array_equal(got[q], ref[q])
    assert got["D"].dtype == np.float32 and got["d"].dtype == np.float32
    # tolerance: float32 rounding of a float64 value, 2e-6 relative to |D|
    np.testing.assert_allclose(got["d"], ref["d"], rtol=2e-6, atol=0)
    scale = np.maximum(np.abs(ref["D"]), ref["d"][:, None])
    assert np.all(np.abs(got["D"] - ref["D"]) <= 2e-6 * scale)


@pytest.mark.parametrize("order", ["zyx", "xyz", "yzx", "random"])
def test_layout_follows_atom_order(gpu, order):
    """The dense cell layout follows the previous build's atom order (an
    internal choice: outputs must not change). A 9x11x13 fcc lattice in
    C order (z fastest), Fortran order (x fastest), y fastest and random
    order, built four times in a row on the same plan, bit for bit against
    the oracle each time."""
    a, reps = 3.6, (9, 11, 13)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    grids = np.stack(np.meshgrid(*[np.arange(r) for r in reps], indexing="ij"), -1)
    if order == "xyz":
        grids = grids.transpose(2, 1, 0, 3)
    elif order == "yzx":
        grids = grids.transpose(0, 2, 1, 3)
    pos = ((grids.reshape(-1, 1, 3) + basis[None]) * a).reshape(-1, 3)
    rng = np.random.default_rng(17)
    if order == "random":
        pos = pos[rng.permutation(len(pos))]
    pos = pos + rng.normal(0.0, 0.05, pos.shape)
    cell = np.diag(np.asarray(reps, float) * a)
    for rep in range(4):
        _check(pos, cell, True, 4.2, msg=f"{order} build {rep}")


@pytest.mark.parametrize("cut", ["global", "per_atom", "per_type", "f32", "matrix"])
def test_five_chunk_fill_window(gpu, cut):
    """A density at which a home cell has 257-320 candidates (fcc 14^3,
    a = 3.615, rc 5: 10 cells a side, ~296 candidates), so the fill takes
    its 5-chunk window: uniform, per-atom and per-type cutoffs (dense and
    hashed grids) bit for bit, float32 output to its tolerance, and the dense
    neighbour matrix (rows short of and above the largest count)."""
    a, reps = 3.615, 14
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(*[np.arange(reps)] * 3, indexing="ij"), -1).reshape(-1, 1, 3)
    pos = ((g + basis[None]) * a).reshape(-1, 3)
    rng = np.random.default_rng(23)
    pos = pos + rng.normal(0.0, 0.1, pos.shape)
    cell = reps * a * np.eye(3)
    nat = len(pos)
    assert 256 < 27 * nat / 10 ** 3 <= 320
    if cut == "global":
        _check(pos, cell, True, 5.0, backends=(0, 1))
    elif cut == "per_atom":
        _check(pos, cell, True, rng.uniform(2.2, 2.5, nat), backends=(0, 1))
    elif cut == "per_type":
        types = rng.integers(0, 2, nat).astype(np.int32)
        _check(pos, cell, True, np.array([[4.0, 4.6], [4.6, 5.0]]), numbers=types,
               backends=(0, 1))
    elif cut == "matrix":
        from conftest import bits
        from test_gpu_matrix import _engine_matrix
        inv = np.ascontiguousarray(np.linalg.inv(cell.T))
        pb = np.ones(3, bool)
        kmax = int(oracle.neighbour_matrix(np.zeros(3), cell, inv, pb, pos, 5.0, 1)[2].max())
        for K in (kmax - 2, kmax + 3):
            want = oracle.neighbour_matrix(np.zeros(3), cell, inv, pb, pos, 5.0, K)
            for backend in (0, 1):
                got = _engine_matrix(pos, cell, inv, pb, 5.0, None, K, backend)
                assert np.array_equal(got[0], want[0]) and np.array_equal(got[2], want[2])
                assert np.array_equal(bits(got[1]), bits(want[1]))
                assert got[3] == want[3]
    else:
        ref = _oracle("ijdDS", pos, cell, True, 5.0)
        got = _engine("ijdDS", pos, cell, True, 5.0, f32=True)
        for q in "ijS":
            assert np.array_equal(got[q], ref[q])
        np.testing.assert_allclose(got["d"], ref["d"], rtol=2e-6, atol=0)
        scale = np.maximum(np.abs(ref["D"]), ref["d"][:, None])
        assert np.all(np.abs(got["D"] - ref["D"]) <= 2e-6 * scale)


@pytest.mark.parametrize("seed", range(24))
def test_build_sequence_fuzz(gpu, seed):
    """Random systems built twice each on the same plan (the second build
    uses the cell layout the first one's atom order selected), bit for bit
    against the oracle, in both grid backends: edge lengths from well below
    to many times the cutoff (one or several periodic images, stencil reach
    beyond the grid), orthorhombic and triclinic cells, mixed periodicity,
    atoms in random order or sorted along a random axis order, and every
    cutoff form (uniform, per-atom radii, symmetric and asymmetric per-type
    tables)."""
    rng = np.random.default_rng(5000 + seed)
    n = int(rng.integers(1, 2500))
    rc = float(rng.uniform(0.8, 3.0))
    edges = rng.uniform(0.6, 14.0, 3) * rc / 2.0
    cell = np.diag(edges)
    if seed % 3 == 1:
        cell = cell + rng.uniform(-0.2, 0.2, (3, 3)) * edges.min()
    pbc = rng.integers(0, 2, 3).astype(bool)
    pos = rng.uniform(-0.2, 1.2, (n, 3)) @ cell
    if seed % 2 == 0:
        f, m, s = rng.permutation(3)
        pos = pos[np.lexsort((pos[:, f], np.floor(pos[:, m] / rc), np.floor(pos[:, s] / rc)))]
    mode = seed % 4
    numbers = None
    if mode == 0:
        cutoff = rc
    elif mode == 1:
        cutoff = rng.uniform(0.3, 0.5, n) * rc
    else:
        numbers = rng.integers(0, 3, n).astype(np.int32)
        t = rng.uniform(0.5, 1.0, (3, 3)) * rc
        cutoff = (t + t.T) / 2 if mode == 2 else t
    for rep in range(2):
        _check(pos, cell, pbc, cutoff, numbers=numbers, backends=(0, 1),
               msg=f"seed={seed} build {rep}")


@pytest.mark.parametrize("case", ["atoms", "cells"])
def test_small_binning_limits(gpu, case):
    """Single dense systems of up to 4096 atoms and 4096 cells are binned by
    one block (grid, scan, member lists and stable ranks in LDS); larger
    ones by the multi-kernel binning. Sizes on both sides of each limit,
    built one after another on the same plan (the paths hand the plan's
    parameters and order sample to each other), bit for bit against the
    oracle."""
    rng = np.random.default_rng(71 if case == "atoms" else 72)
    if case == "atoms":
        # 12 cells a side (1728 cells), 4095 / 4096 / 4097 atoms
        L, rc = 12.5, 1.0
        sizes = [(4096, (L, L, L)), (4097, (L, L, L)), (4095, (L, L, L)), (4096, (L, L, L))]
    else:
        # 16 x 16 x 16 = 4096 cells against 17 x 16 x 16 = 4352
        rc = 1.0
        sizes = [(2000, (16.5, 16.5, 16.5)), (2000, (17.5, 16.5, 16.5)),
                 (2000, (16.5, 16.5, 16.5))]
    for k, (n, edges) in enumerate(sizes):
        cell = np.diag(edges)
        pos = rng.uniform(0.0, 1.0, (n, 3)) @ cell
        _check(pos, cell, True, rc, backends=(-1, 0), msg=f"{case} build {k}: {n} atoms")


@pytest.mark.parametrize("order", ["zyx", "xyz", "random"])
def test_small_binning_layout(gpu, order):
    """The one-block binning takes the layout the previous build's order
    sample selects and samples the input for the next build: a 6x7x8 fcc
    lattice (1344 atoms) in C order, Fortran order and random order, built
    three times in a row, bit for bit against the oracle each time."""
    a, reps = 3.6, (6, 7, 8)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    grids = np.stack(np.meshgrid(*[np.arange(r) for r in reps], indexing="ij"), -1)
    if order == "xyz":
        grids = grids.transpose(2, 1, 0, 3)
    pos = ((grids.reshape(-1, 1, 3) + basis[None]) * a).reshape(-1, 3)
    rng = np.random.default_rng(19)
    if order == "random":
        pos = pos[rng.permutation(len(pos))]
    pos = pos + rng.normal(0.0, 0.05, pos.shape)
    cell = np.diag(np.asarray(reps, float) * a)
    for rep in range(3):
        _check(pos, cell, True, 4.2, msg=f"{order} build {rep}")


@pytest.mark.parametrize("pbc", [True, False])
def test_small_binning_one_full_cell(gpu, pbc):
    """600 atoms in a 4 A box at rc 3: a single cell holds every atom (the
    stable rank compares against 600 members), and with pbc its 27 images
    are the stencil."""
    rng = np.random.default_rng(73)
    cell = 4.0 * np.eye(3)
    pos = rng.uniform(0.0, 4.0, (600, 3))
    _check(pos, cell, pbc, 3.0, backends=(-1, 0, 1))
