"""first_neighbours, get_jump_indicies and triplet_list on the GPU against
the compiled reference's known answers (tests/test_extension.py:277-337)
and the oracle's restatements on real pair lists, host and device inputs."""

import numpy as np
import pytest

import oracle
from test_oracle import INDEX

pytestmark = pytest.mark.gpu


def test_known_answers(gpu):
    from matscipy_neighbours_amd import first_neighbours, get_jump_indicies, triplet_list
    post = INDEX["post"]
    for c in post["first_neighbours"]:
        assert first_neighbours(c["n"], np.array(c["i"], np.int32)).tolist() == c["seed"]
    for c in post["get_jump_indicies"]:
        assert get_jump_indicies(np.array(c["sorted"], np.int32)).tolist() == c["out"]
    t0, t1 = post["triplet_list"]
    ij, ik = triplet_list(np.array(t0["first_i"], np.int32))
    assert ij.tolist() == t0["ij"] and ik.tolist() == t0["ik"]
    ij, ik = triplet_list(np.array(t1["first_i"], np.int32), np.array(t1["absdist"]),
                          t1["cutoff"])
    assert ij.tolist() == t1["ij"] and ik.tolist() == t1["ik"]


def test_on_pair_lists(gpu):
    """Real pair lists (atoms with no neighbours included), host and device."""
    import torch

    from matscipy_neighbours_amd import (first_neighbours, get_jump_indicies,
                                         neighbour_list, triplet_list)
    rng = np.random.default_rng(12)
    pos = rng.uniform(0, 14, (400, 3))
    pos[:40] += 100.0  # isolated atoms far outside (no neighbours)
    for nonperiodic in (False, True):
        i, d = neighbour_list("id", positions=pos, cell=14 * np.eye(3), pbc=not nonperiodic,
                              cutoff=2.0)
        seed = first_neighbours(len(pos), i)
        want = oracle.first_neighbours(len(pos), i)
        assert np.array_equal(seed, want)
        # get_jump_indicies takes values 0 .. n-1 (it indexes seed[value]):
        # relabel the atoms that have pairs densely
        dense = np.unique(i, return_inverse=True)[1].astype(np.int32)
        assert np.array_equal(get_jump_indicies(dense), oracle.get_jump_indicies(dense))
        for args in ((seed,), (seed, d, 1.5)):
            got = triplet_list(*args)
            ref = oracle.triplet_list(*args)
            assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        # device in, device out
        it = torch.from_numpy(i).to("cuda:0")
        sd = first_neighbours(len(pos), it)
        assert isinstance(sd, torch.Tensor) and np.array_equal(sd.cpu().numpy(), want)
        ij, ik = triplet_list(sd, torch.from_numpy(d).to("cuda:0"), 1.5)
        ref = oracle.triplet_list(seed, d, 1.5)
        assert np.array_equal(ij.cpu().numpy(), ref[0]) and np.array_equal(ik.cpu().numpy(),
                                                                           ref[1])


def test_empty_and_edge_cases(gpu):
    from matscipy_neighbours_amd import first_neighbours, get_jump_indicies, triplet_list
    assert first_neighbours(4, np.zeros(0, np.int32)).tolist() == [0] * 5
    assert get_jump_indicies(np.zeros(0, np.int32)).tolist() == \
        oracle.get_jump_indicies(np.zeros(0, np.int32)).tolist()
    ij, ik = triplet_list(np.array([0, 1, 1, 3], np.int32))
    assert ij.tolist() == [1, 2] and ik.tolist() == [2, 1]
    d = np.array([0.5, 1.0, np.nan])  # NaN is kept (not >= cutoff)
    ij, ik = triplet_list(np.array([0, 3], np.int32), d, 0.9)
    ref = oracle.triplet_list(np.array([0, 3], np.int32), d, 0.9)
    assert ij.tolist() == ref[0].tolist() and ik.tolist() == ref[1].tolist()
    # rows before the first atom with pairs (-1), empty rows in the middle
    # and at the end, with and without a cutoff
    seed = np.array([-1, -1, 0, 3, 3, 5, 5], np.int32)
    d = np.array([0.1, 0.2, 2.0, 0.3, 0.4])
    for args in ((seed,), (seed, d, 1.0)):
        got, ref = triplet_list(*args), oracle.triplet_list(*args)
        assert got[0].tolist() == ref[0].tolist() and got[1].tolist() == ref[1].tolist()
    assert oracle.triplet_list(seed, d, 1.0)[0].tolist() == [0, 1, 3, 4]


@pytest.mark.parametrize("with_cutoff", [False, True])
def test_triplet_row_lengths(gpu, with_cutoff):
    """Rows of 0 to 200 pairs, whose triplet indices step by 64 / (m - 1)
    and 64 mod (m - 1) per lane pass (m - 1 below, at and above 64), over
    331 rows (64 rows per wave, the last wave partial), with and without a
    distance cutoff, against the oracle; host and device input."""
    import torch

    from matscipy_neighbours_amd import triplet_list
    rng = np.random.default_rng(31)
    lengths = np.concatenate([[0, 1, 2, 3, 63, 64, 65, 66, 127, 128, 129, 200, 0, 5],
                              rng.integers(0, 40, 317)])
    seed = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int32)
    args = (seed,)
    if with_cutoff:
        d = rng.uniform(0.0, 2.0, int(seed[-1]))
        args = (seed, d, 1.3)
    want = oracle.triplet_list(*args)
    got = triplet_list(*args)
    assert len(want[0]) > 100_000
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    dev = tuple(torch.from_numpy(np.asarray(a)).to("cuda:0") if isinstance(a, np.ndarray)
                else a for a in args)
    ij, ik = triplet_list(*dev)
    assert np.array_equal(ij.cpu().numpy(), want[0]) and np.array_equal(ik.cpu().numpy(),
                                                                        want[1])


def test_first_neighbours_alignment_and_tails(gpu):
    """first_neighbours reads four pairs per thread where the list is 16-byte
    aligned and one per thread for the rest: device slices starting 0 to 3
    elements into the list (aligned and not) and lengths that leave 0 to 3
    pairs over, against the oracle."""
    import torch

    from matscipy_neighbours_amd import first_neighbours
    rng = np.random.default_rng(17)
    i = np.sort(rng.integers(3, 5000, 40_003)).astype(np.int32)
    it = torch.from_numpy(i).to("cuda:0")
    for start in range(4):
        for stop in (len(i), len(i) - 1, len(i) - 2, len(i) - 3, start + 5, start + 1):
            got = first_neighbours(5003, it[start:stop])
            want = oracle.first_neighbours(5003, i[start:stop])
            assert np.array_equal(got.cpu().numpy(), want), (start, stop)
