"""Index logic of the post-processing kernels (csrc/mnl_post.hip), restated
on the host with every memory access bounds-checked, against the oracle's
first_neighbours / triplet_list (src/libneighbours/first_neighbours.cc:15-40,
triplet_list.cc:15-34).

The lists have leading atoms without pairs, so their seeds are -1, and
trailing atoms without pairs (seeds == nn): the shapes on which a kernel
that took a row's first index at face value would read absdist[-1] and
store vidx[-1]. The restatements follow the kernels statement for statement
(per thread / per wave), and the accessor classes raise on any index outside
the buffer, so a lost guard fails here on the CPU.
"""

import numpy as np
import pytest

import oracle


class Checked:
    """A 1-D buffer whose reads and writes raise outside [0, len)."""

    def __init__(self, arr, name):
        self.a = arr
        self.name = name

    def _chk(self, k):
        if not (0 <= k < len(self.a)):
            raise IndexError(f"{self.name}[{k}] outside [0, {len(self.a)})")

    def __getitem__(self, k):
        self._chk(k)
        return self.a[k]

    def __setitem__(self, k, v):
        self._chk(k)
        self.a[k] = v


def seed_at(k, ik, ip, nn, n, seed):
    """mnl_post.hip seed_at."""
    if k == 0:
        if 0 <= ik <= n:
            seed[ik] = 0
    else:
        hi = ik if ik < n else n
        for l in range(max(ip + 1, 0), hi + 1):
            seed[l] = k
    if k == nn - 1:
        for l in range(max(ik + 1, 0), n + 1):
            seed[l] = nn


def first_neighbours_host(n, i):
    """launch_first_neighbours: fill with -1 (0 for an empty list), then one
    seed_at per pair (the 4-wide kernel calls the same function)."""
    nn = len(i)
    seed = Checked(np.full(n + 1, -1 if nn > 0 else 0, np.int64), "seed")
    ic = Checked(i, "i")
    for k in range(nn):
        seed_at(k, int(ic[k]), int(ic[k - 1]) if k > 0 else 0, nn, n, seed)
    return seed.a


def triplet_rows_host(first_i, nrows, nn, absdist, cutoff):
    """k_triplet_rows: one wave per 64 consecutive rows."""
    fi = Checked(first_i, "first_i")
    ad = Checked(absdist, "absdist")
    vidx = Checked(np.full(max(nn, 1), -9, np.int64), "vidx")
    m = Checked(np.zeros(nrows, np.int64), "m")
    vstart = Checked(np.zeros(nrows, np.int64), "vstart")
    for r0 in range(0, nrows, 64):
        nr = min(64, nrows - r0)
        b = [int(fi[r0 + l]) if l < nr else -1 for l in range(64)]
        hi = min(int(fi[r0 + nr]), nn)
        lo = min([x for x in b if x >= 0], default=2 ** 31 - 1)
        gp = [-1] * 64
        run = 0
        for k0 in range(lo, hi, 64):
            keep = [k0 + l < hi and not (ad[k0 + l] >= cutoff) for l in range(64)]
            pre = np.concatenate([[0], np.cumsum(keep)])
            for l in range(64):
                if keep[l]:
                    vidx[lo + run + int(pre[l])] = k0 + l
            for l in range(64):
                if k0 <= b[l] < k0 + 64:
                    gp[l] = run + int(pre[b[l] - k0])
            run += int(pre[64])
        for l in range(64):
            if b[l] >= 0 and gp[l] < 0:
                gp[l] = run
        for l in range(nr):
            gnext = run if l == nr - 1 else gp[l + 1]
            cnt = gnext - gp[l] if b[l] >= 0 else 0
            m[r0 + l] = cnt
            vstart[r0 + l] = lo + gp[l] if b[l] >= 0 else 0
    return m.a, vstart.a, vidx.a


def triplet_rows_all_host(first_i, nrows, nn):
    """k_triplet_rows_all: one thread per row."""
    fi = Checked(first_i, "first_i")
    m = np.zeros(nrows, np.int64)
    for r in range(nrows):
        b = int(fi[r])
        e = min(int(fi[r + 1]), nn)
        m[r] = e - b if (b >= 0 and e > b) else 0
    return m


def triplets_from_rows(m, vbase, vidx):
    """k_triplet_fill's order: per row, ij outer, ik inner, both ascending."""
    ij, ik = [], []
    for r in range(len(m)):
        mr = int(m[r])
        for a in range(mr):
            for c in range(mr):
                if c != a:
                    ij.append(vidx[vbase[r] + a] if vidx is not None else vbase[r] + a)
                    ik.append(vidx[vbase[r] + c] if vidx is not None else vbase[r] + c)
    return np.array(ij, np.int64), np.array(ik, np.int64)


def _list(rng, n, lead, tail):
    """An i-sorted list over n atoms whose first `lead` and last `tail`
    atoms have no pairs, other atoms 0-7 pairs."""
    counts = rng.integers(0, 8, n)
    counts[:lead] = 0
    if tail:
        counts[n - tail:] = 0
    i = np.repeat(np.arange(n, dtype=np.int32), counts)
    d = rng.uniform(0.5, 3.0, len(i))
    return i, d


@pytest.mark.parametrize("seed,n,lead,tail", [(0, 100, 40, 0), (1, 130, 40, 7), (2, 70, 0, 20),
                                              (3, 200, 64, 64), (4, 65, 63, 1), (5, 10, 9, 0)])
def test_seeds_and_triplet_rows_bounds(seed, n, lead, tail):
    rng = np.random.default_rng(seed)
    i, d = _list(rng, n, lead, tail)
    nn = len(i)
    seed_arr = first_neighbours_host(n, i)
    want = oracle.first_neighbours(n, i)
    assert np.array_equal(seed_arr, want)
    assert np.all(seed_arr[:lead] == -1) if lead and nn else True
    first_i = np.ascontiguousarray(want, np.int64)
    nrows = n  # n_first - 1 rows
    for cutoff in (None, 2.0):
        if cutoff is None:
            m = triplet_rows_all_host(first_i, nrows, nn)
            got = triplets_from_rows(m, first_i, None)
            ref = oracle.triplet_list(first_i.astype(np.int32))
        else:
            m, vstart, vidx = triplet_rows_host(first_i, nrows, nn, d, cutoff)
            got = triplets_from_rows(m, vstart, vidx)
            ref = oracle.triplet_list(first_i.astype(np.int32), d, cutoff)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), cutoff


def test_unguarded_row_start_reads_before_the_buffer():
    """The guard matters: a row bound taken from a -1 seed without the
    b >= 0 test reaches index -1 (the read and store that the guard
    prevents)."""
    rng = np.random.default_rng(7)
    i, d = _list(rng, 100, 40, 0)
    first_i = oracle.first_neighbours(100, i).astype(np.int64)
    ad = Checked(d, "absdist")
    b, e = int(first_i[39]), min(int(first_i[40]), len(i))
    assert b == -1 and e == 0
    with pytest.raises(IndexError):
        for k in range(b, e):  # a loop over [first_i[r], first_i[r+1]) with no guard
            ad[k]
