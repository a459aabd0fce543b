"""The reference's compatibility tests (tests/test_matscipy_compat.py), which
need ASE, restated with explicit arrays: lattices built by hand, formic
acid from standard bond lengths, and an ``atoms``-like object for the
``atoms=`` argument. Expected values are the reference tests' own."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Atoms:
    """The attributes neighbour_list reads from an ASE Atoms object."""

    def __init__(self, positions, cell, pbc, numbers):
        self.positions = np.asarray(positions, float)
        self.cell = np.asarray(cell, float)
        self.pbc = np.asarray(pbc, bool)
        self.numbers = np.asarray(numbers, np.int32)


def _lattice(basis_frac, cell, reps):
    basis = np.asarray(basis_frac, float)
    g = np.stack(np.meshgrid(*[np.arange(r) for r in reps], indexing="ij"), -1).reshape(-1, 1, 3)
    frac = ((g + basis[None]) / np.asarray(reps, float)).reshape(-1, 3)
    big = np.asarray(cell, float) * np.asarray(reps, float)[:, None]
    return frac @ big, big


FCC = [[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]
DIAMOND = FCC + [[0.25, 0.25, 0.25], [0.75, 0.75, 0.25], [0.75, 0.25, 0.75], [0.25, 0.75, 0.75]]


def test_fcc_primitive_coordination_and_distance(gpu):
    from matscipy_neighbours_amd import neighbour_list
    a = 4.05
    cell = 0.5 * a * np.array([[0.0, 1, 1], [1, 0, 1], [1, 1, 0]])
    atoms = _Atoms(np.zeros((1, 3)), cell, True, [13])
    i, j, d = neighbour_list("ijd", atoms, 3.1)
    assert (np.bincount(i) == [12]).all()
    np.testing.assert_allclose(d, a / np.sqrt(2), atol=1e-6)


@pytest.mark.parametrize("n", [1, 2, 3])
def test_graphite_coordination_is_three(gpu, n):
    from matscipy_neighbours_amd import coordination
    a, c = 2.5, 10.0
    hexcell = np.array([[a, 0, 0], [-a / 2, a * np.sqrt(3) / 2, 0], [0, 0, c]])
    basis = [[0, 0, 0], [1 / 3, 2 / 3, 0], [0, 0, 0.5], [2 / 3, 1 / 3, 0.5]]
    pos, big = _lattice(basis, hexcell, (n, n, 1))
    assert (coordination(_Atoms(pos, big, True, [6] * len(pos)), 1.85) == 3).all()


def _rattled(basis, a, reps, sigma, seed):
    pos, big = _lattice(basis, a * np.eye(3), reps)
    pos = pos + np.random.default_rng(seed).normal(0.0, sigma, pos.shape)
    return pos, big


def test_shift_reconstructs_distance_vector(gpu):
    from matscipy_neighbours_amd import neighbour_list
    pos, cell = _rattled(FCC, 3.6, (1, 1, 1), 0.1, 12)
    i, j, D, S = neighbour_list("ijDS", _Atoms(pos, cell, True, [29] * 4), 3.0)
    np.testing.assert_allclose(D, pos[j] - pos[i] + S.dot(cell), atol=1e-10)


def test_distance_vector_matches_mic(gpu):
    from matscipy_neighbours_amd import mic, neighbour_list
    pos, cell = _rattled(FCC, 3.6, (2, 2, 2), 0.1, 3)
    i, j, D, d = neighbour_list("ijDd", _Atoms(pos, cell, True, [29] * len(pos)), 3.0)
    np.testing.assert_allclose(D, mic(pos[j] - pos[i], cell), atol=1e-10)
    np.testing.assert_allclose(d, np.linalg.norm(D, axis=1), atol=1e-12)


def test_pair_list_is_symmetric(gpu):
    from matscipy_neighbours_amd import neighbour_list
    pos, cell = _lattice(DIAMOND, 5.43 * np.eye(3), (1, 1, 1))
    i, j = neighbour_list("ij", _Atoms(pos, cell, True, [14] * 8), 2.6)
    assert (np.bincount(i) == np.bincount(j)).all()


def _formic_acid():
    """HCOOH in the order O, C, O, H(on O), H(on C), planar, standard bond
    lengths (C-OH 1.34, C=O 1.20, O-H 0.97, C-H 1.10 A), centred in vacuum."""
    C = np.array([0.0, 0.0, 0.0])
    ang = np.deg2rad
    O1 = C + 1.34 * np.array([np.cos(ang(-120)), np.sin(ang(-120)), 0])
    O2 = C + 1.20 * np.array([1.0, 0.0, 0.0])
    H2 = C + 1.10 * np.array([np.cos(ang(120)), np.sin(ang(120)), 0])
    H1 = O1 + 0.97 * np.array([np.cos(ang(-120)), np.sin(ang(-120)), 0])
    pos = np.array([O1, C, O2, H1, H2])
    pos = pos - pos.min(axis=0) + 5.0
    cell = np.diag(pos.max(axis=0) + 5.0)
    return _Atoms(pos, cell, False, [8, 6, 8, 1, 1])


def test_element_pair_dict_cutoffs(gpu):
    from matscipy_neighbours_amd import neighbour_list
    a = _formic_acid()
    assert (np.bincount(neighbour_list("i", a, 1.85)) == [2, 3, 1, 1, 1]).all()
    assert (np.bincount(neighbour_list("i", a, {(1, 6): 1.2})) == [0, 1, 0, 0, 1]).all()
    assert (np.bincount(neighbour_list("i", a, {("C", "O"): 1.4})) == [1, 2, 1]).all()
    assert (np.bincount(neighbour_list("i", a, {("H", "C"): 1.2, (6, 8): 1.4})) ==
            [1, 3, 1, 0, 1]).all()


def test_per_atom_radii_cutoff(gpu):
    from matscipy_neighbours_amd import neighbour_list
    a = _formic_acid()
    i_pa = neighbour_list("i", a, np.full(5, 0.9))
    i_global = neighbour_list("i", a, 1.8)
    assert (np.bincount(i_pa, minlength=5) == np.bincount(i_global, minlength=5)).all()


def test_atoms_outside_box_are_wrapped(gpu):
    from matscipy_neighbours_amd import neighbour_list
    pos, cell = _lattice(FCC, 3.6 * np.eye(3), (1, 1, 1))
    ref_i = neighbour_list("i", _Atoms(pos, cell, True, [29] * 4), 3.0)
    moved = pos.copy()
    moved[0] += cell[0] + 2 * cell[1]
    out_i = neighbour_list("i", _Atoms(moved, cell, True, [29] * 4), 3.0)
    assert (np.bincount(ref_i) == np.bincount(out_i)).all()


def test_supercell_coordination_consistency(gpu):
    from matscipy_neighbours_amd import neighbour_list
    pos, cell = _lattice(DIAMOND, 5.43 * np.eye(3), (1, 1, 1))
    c1 = np.bincount(neighbour_list("i", _Atoms(pos, cell, True, [14] * 8), 2.6))
    pos2, cell2 = _lattice(DIAMOND, 5.43 * np.eye(3), (2, 1, 2))
    c2 = np.bincount(neighbour_list("i", _Atoms(pos2, cell2, True, [14] * len(pos2)), 2.6))
    assert set(np.unique(c1)) == set(np.unique(c2))
    assert (c2 == c1[0]).all()


def test_shrink_wrapped_positions_only(gpu):
    from matscipy_neighbours_amd import neighbour_list
    r = np.random.default_rng(0).uniform(0, 8, size=(30, 3))
    i, j, D, d = neighbour_list("ijDd", positions=r, cutoff=1.5)
    np.testing.assert_allclose(D, r[j] - r[i], atol=1e-12)
    np.testing.assert_allclose(d, np.linalg.norm(D, axis=1), atol=1e-12)
    assert (np.bincount(i) == np.bincount(j)).all()


def test_atoms_and_explicit_arrays_are_exclusive(gpu):
    from matscipy_neighbours_amd import neighbour_list
    a = _formic_acid()
    with pytest.raises(ValueError):
        neighbour_list("i", a, 1.85, positions=a.positions)


def test_coordination_matches_oracle_back_to_back(gpu):
    """coordination leaves the pair total unread (nothing waits for the
    device): two calls in a row on device positions, then the host copies of
    both results, each equal to the oracle's per-atom pair counts; and the
    host-array form."""
    import torch

    import oracle
    from matscipy_neighbours_amd import coordination
    rng = np.random.default_rng(44)
    cell = np.array([[11.0, 0.0, 0.0], [1.5, 12.0, 0.0], [-0.7, 0.9, 10.5]])
    inv = np.ascontiguousarray(np.linalg.inv(cell.T))
    pos = [rng.uniform(-0.1, 1.1, (1800, 3)) @ cell for _ in range(2)]
    got = [coordination(positions=torch.from_numpy(p).to("cuda:0"), cell=cell,
                        pbc=(True, True, False), cutoff=2.7) for p in pos]
    for p, g in zip(pos, got):
        ref = oracle.neighbour_list("i", np.zeros(3), cell, inv, np.array([True, True, False]),
                                    p, 2.7)["i"]
        want = np.bincount(ref, minlength=len(p))
        assert np.array_equal(g.cpu().numpy(), want)
        assert np.array_equal(coordination(positions=p, cell=cell, pbc=(True, True, False),
                                           cutoff=2.7), want)


def test_repeated_device_builds_with_growing_totals(gpu):
    """Device builds that repeat the plan's previous one (same atom count and
    cutoff) queue the fill with a capacity from the previous pair total: a
    second build with about the same total fits, a third in a box half the
    size (eight times the pairs) overflows it and fills again at the exact
    size. Every result equals the oracle bit for bit, and the returned rows
    are exactly the pair total."""
    import torch

    import oracle
    from conftest import assert_same_pairs
    from matscipy_neighbours_amd import _lib, neighbour_list
    if hasattr(_lib.plan_for(0), "last_build"):
        del _lib.plan_for(0).last_build
    rng = np.random.default_rng(45)
    frac = rng.uniform(0.0, 1.0, (4000, 3))
    for k, edge in enumerate((24.0, 24.0, 12.0)):
        cell = edge * np.eye(3)
        pos = frac @ cell + (rng.normal(0.0, 0.01, frac.shape) if k == 1 else 0.0)
        got = neighbour_list("ijdDS", positions=torch.from_numpy(pos).to("cuda:0"), cell=cell,
                             pbc=True, cutoff=3.0)
        got = dict(zip("ijdDS", (t.cpu().numpy() for t in got)))
        inv = np.ascontiguousarray(np.linalg.inv(cell.T))
        ref = oracle.neighbour_list("ijdDS", np.zeros(3), cell, inv, np.ones(3, bool), pos, 3.0)
        assert_same_pairs(got, ref, msg=f"build {k}")
        assert len(got["i"]) == len(ref["i"])
