"""mnl oracle: CPU checker for the HIP neighbour-list engine.

TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this package; the product path never does.

Two CPU libraries are reachable from here:

* ``libmnl_oracle.so`` (oracle/mnl_oracle.c) - the restatement of the
  reference CPU path, operation for operation. It is pinned against the
  golden fixtures written from the compiled reference, against the
  reference's own known-answer tests, and against an independent brute-force
  search (tests/test_oracle.py).
* ``_ref/libref_neighbours.so`` - the reference CPU core compiled from its
  own sources by oracle/Makefile (only where the reference checkout exists
  when building). Used to write the fixtures, to cross-check the
  restatement, and as the CPU baseline in bench.py.

Argument handling (scalar / per-atom / per-type cutoffs) restates
language_bindings/python/bind_py_neighbours.cc:130-195.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(_HERE, "libmnl_oracle.so")
REF_LIB = os.path.join(_HERE, "_ref", "libref_neighbours.so")

QUANTITY_BITS = {"i": 1, "j": 2, "D": 4, "d": 8, "S": 16}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lp = ctypes.POINTER(ctypes.c_int64)


class _OracleResult(ctypes.Structure):
    _fields_ = [("npairs", ctypes.c_int64), ("first", _ip), ("secnd", _ip),
                ("distvec", _dp), ("absdist", _dp), ("shift", _ip),
                ("counts", _ip), ("sparse", ctypes.c_int)]


class _RefResult(ctypes.Structure):
    _fields_ = [("npairs", ctypes.c_int64), ("first", _ip), ("secnd", _ip),
                ("distvec", _dp), ("absdist", _dp), ("shift", _ip)]


_oracle = None
_ref = None


def _load_oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"oracle library not built: {ORACLE_LIB} "
                               "(run `make -C oracle`)")
        lib = ctypes.CDLL(ORACLE_LIB)
        lib.oracle_neighbour_list.restype = ctypes.c_int
        lib.oracle_neighbour_list.argtypes = [
            ctypes.c_int, _dp, _dp, _dp, _ip, ctypes.c_int32, _dp, ctypes.c_double,
            _dp, _dp, ctypes.c_int32, _ip, ctypes.c_int,
            ctypes.POINTER(_OracleResult)]
        lib.oracle_free.argtypes = [ctypes.POINTER(_OracleResult)]
        lib.oracle_last_error.restype = ctypes.c_char_p
        lib.oracle_grid_geometry.argtypes = [_dp, ctypes.c_double, _ip, _ip, _dp, _dp]
        lib.oracle_first_neighbours.argtypes = [ctypes.c_int32, ctypes.c_int32, _ip, _ip]
        _oracle = lib
    return _oracle


def ref_available():
    return os.path.exists(REF_LIB)


def _load_ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"compiled reference not built: {REF_LIB}")
        lib = ctypes.CDLL(REF_LIB)
        lib.ref_neighbour_list.restype = ctypes.c_int
        lib.ref_neighbour_list.argtypes = [
            ctypes.c_int, _dp, _dp, _dp, _ip, ctypes.c_int32, _dp, ctypes.c_double,
            _dp, _dp, ctypes.c_int32, _ip, ctypes.POINTER(_RefResult)]
        lib.ref_free.argtypes = [ctypes.POINTER(_RefResult)]
        lib.ref_last_error.restype = ctypes.c_char_p
        lib.ref_time_neighbour_list.restype = ctypes.c_int
        lib.ref_time_neighbour_list.argtypes = [
            ctypes.c_int, _dp, _dp, _dp, _ip, ctypes.c_int32, _dp, ctypes.c_double,
            ctypes.c_int32, _dp, _lp]
        lib.ref_max_threads.restype = ctypes.c_int
        lib.ref_set_threads.argtypes = [ctypes.c_int]
        lib.ref_first_neighbours.argtypes = [ctypes.c_int32, ctypes.c_int32, _ip, _ip]
        lib.ref_get_jump_indicies.restype = ctypes.c_int32
        lib.ref_get_jump_indicies.argtypes = [ctypes.c_int32, _ip, _ip, ctypes.c_int32]
        lib.ref_triplet_list.restype = ctypes.c_int64
        lib.ref_triplet_list.argtypes = [ctypes.c_int32, _ip, _dp, ctypes.c_double,
                                         _ip, _ip, ctypes.c_int64]
        lib.ref_neighbour_matrix.restype = ctypes.c_int
        lib.ref_neighbour_matrix.argtypes = [
            _dp, _dp, _dp, _ip, ctypes.c_int32, _dp, ctypes.c_double, _dp, _dp,
            ctypes.c_int32, _ip, ctypes.c_int32, _ip, _dp, _ip, _ip]
        _ref = lib
    return _ref


def _ptr(a, kind):
    if a is None:
        return None
    return a.ctypes.data_as(_dp if kind == "d" else _ip)


def quantity_flags(quantities):
    """bind_py_neighbours.cc:183-195: unknown characters raise ValueError."""
    flags = 0
    for q in quantities:
        if q not in QUANTITY_BITS:
            raise ValueError("Unsupported quantity specified.")
        flags |= QUANTITY_BITS[q]
    return flags


def resolve_cutoff(cutoff, nat):
    """bind_py_neighbours.cc:130-173: a Python float is the scalar cutoff; a
    1-d array is per-atom radii (binning cutoff = max 2r); a 2-d square array is
    the per-type matrix (binning cutoff = max, squares stored).
    Returns (cutoff, per_atom or None, per_type_sq or None, ncutoffs)."""
    if isinstance(cutoff, float):
        return float(cutoff), None, None, 0
    a = np.ascontiguousarray(np.asarray(cutoff, dtype=np.float64))
    if a.ndim == 1:
        if a.shape[0] != nat:
            raise TypeError("One-dimensional cutoff array must have length that "
                            "corresponds to position array.")
        rc = 0.0
        for v in a:
            rc = max(rc, 2 * float(v))
        return rc, a, None, 0
    if a.ndim == 2:
        if a.shape[0] != a.shape[1]:
            raise TypeError("Two-dimensional cutoff array must be square.")
        rc = 0.0
        flat = a.ravel()
        for v in flat:
            rc = max(rc, float(v))
        return rc, None, np.ascontiguousarray(flat * flat), int(a.shape[0])
    raise ValueError("cutoff must be a float, a 1-d or a 2-d array")


def _prep(cell_origin, cell, inv_cell, pbc, positions, types):
    origin = np.ascontiguousarray(cell_origin, dtype=np.float64).reshape(3)
    cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
    inv_cell = np.ascontiguousarray(inv_cell, dtype=np.float64).reshape(9)
    pbc = np.ascontiguousarray(np.broadcast_to(np.asarray(pbc, dtype=bool), (3,)),
                               dtype=np.int32)
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    ty = None if types is None else np.ascontiguousarray(types, dtype=np.int32)
    return origin, cell, inv_cell, pbc, pos, ty


def _copy(ptr, n, dtype, shape):
    if not ptr or n == 0:
        return np.zeros(shape, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True).reshape(shape)


def neighbour_list(quantities, cell_origin, cell, inv_cell, pbc, positions, cutoff,
                   types=None, backend=-1, with_counts=False):
    """Oracle restatement of the reference ``_ext.neighbour_list``.
    Returns a dict {'i','j','D','d','S'} of the requested quantities (plus
    'counts' per atom when with_counts)."""
    lib = _load_oracle()
    origin, cell, inv_cell, pbc, pos, ty = _prep(cell_origin, cell, inv_cell, pbc,
                                                 positions, types)
    nat = pos.shape[0]
    if ty is not None and ty.shape[0] != nat:
        raise TypeError("Position and type arrays must have identical first dimension.")
    rc, per_atom, per_type_sq, nc = resolve_cutoff(cutoff, nat)
    flags = quantity_flags(quantities)
    res = _OracleResult()
    st = lib.oracle_neighbour_list(flags, _ptr(origin, "d"), _ptr(cell, "d"),
                                   _ptr(inv_cell, "d"), _ptr(pbc, "i"), nat,
                                   _ptr(pos, "d"), rc, _ptr(per_atom, "d"),
                                   _ptr(per_type_sq, "d"), nc, _ptr(ty, "i"),
                                   int(backend), ctypes.byref(res))
    if st != 0:
        msg = lib.oracle_last_error().decode()
        lib.oracle_free(ctypes.byref(res))
        raise RuntimeError(msg)
    n = res.npairs
    out = {}
    try:
        if flags & 1:
            out["i"] = _copy(res.first, n, np.int32, (n,))
        if flags & 2:
            out["j"] = _copy(res.secnd, n, np.int32, (n,))
        if flags & 4:
            out["D"] = _copy(res.distvec, 3 * n, np.float64, (n, 3))
        if flags & 8:
            out["d"] = _copy(res.absdist, n, np.float64, (n,))
        if flags & 16:
            out["S"] = _copy(res.shift, 3 * n, np.int32, (n, 3))
        if with_counts:
            out["counts"] = _copy(res.counts, nat, np.int32, (nat,))
        out["_sparse"] = bool(res.sparse)
    finally:
        lib.oracle_free(ctypes.byref(res))
    return out


def grid_geometry(cell, cutoff):
    """Grid resolution n[3], search range nxyz[3], bins[3,3], len[3]."""
    lib = _load_oracle()
    cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
    n = np.zeros(3, np.int32)
    nxyz = np.zeros(3, np.int32)
    bins = np.zeros(9, np.float64)
    ln = np.zeros(3, np.float64)
    if lib.oracle_grid_geometry(_ptr(cell, "d"), float(cutoff), _ptr(n, "i"),
                                _ptr(nxyz, "i"), _ptr(bins, "d"), _ptr(ln, "d")) != 0:
        raise RuntimeError(lib.oracle_last_error().decode())
    return n, nxyz, bins.reshape(3, 3), ln


def first_neighbours(n, i):
    lib = _load_oracle()
    i = np.ascontiguousarray(i, dtype=np.int32)
    # the C routine indexes seed[i[k]] (as the reference does): values outside
    # [0, n] would write out of bounds
    if i.size and (i.min() < 0 or i.max() > n):
        raise ValueError("first_neighbours: index outside [0, n]")
    seed = np.zeros(n + 1, np.int32)
    lib.oracle_first_neighbours(int(n), int(i.shape[0]), _ptr(i, "i"), _ptr(seed, "i"))
    return seed


def get_jump_indicies(sorted_arr):
    """src/libneighbours/first_neighbours.cc:42-55: the number of distinct
    values (jumps + 1), then first_neighbours of that many rows."""
    s = np.ascontiguousarray(sorted_arr, dtype=np.int32)
    n = 1
    for k in range(len(s) - 1):
        if s[k] != s[k + 1]:
            n += 1
    return first_neighbours(n, s)


def triplet_list(first_i, absdist=None, cutoff=0.0):
    """src/libneighbours/triplet_list.cc:15-34 (pure-Python loops: small
    inputs only): per row, every (ij, ik) with ij != ik, ij outer, ik inner,
    skipping pairs with absdist >= cutoff."""
    f = np.ascontiguousarray(first_i, dtype=np.int64)
    ij_t, ik_t = [], []
    for r in range(len(f) - 1):
        for ij in range(f[r], f[r + 1]):
            for ik in range(f[r], f[r + 1]):
                if ij == ik:
                    continue
                if absdist is not None and (absdist[ij] >= cutoff or absdist[ik] >= cutoff):
                    continue
                ij_t.append(ij)
                ik_t.append(ik)
    return np.array(ij_t, np.int32), np.array(ik_t, np.int32)


def neighbour_matrix(cell_origin, cell, inv_cell, pbc, positions, cutoff, K, types=None):
    """Dense neighbour matrix, restating src/libneighbours/neighbour_list.cc:206-254:
    the pair list (i, j, D) reshaped into rows of K slots in list order, unused
    slots 0, count = row length, overflow if some row is longer than K.
    Returns (idx (nat, K) int32, dist (nat, K, 3) float64, count (nat,) int32,
    overflow bool)."""
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    nat = pos.shape[0]
    idx = np.zeros((nat, K), np.int32)
    dist = np.zeros((nat, K, 3), np.float64)
    cnt = np.zeros(nat, np.int32)
    if nat == 0:
        return idx, dist, cnt, False
    nl = neighbour_list("ijD", cell_origin, cell, inv_cell, pbc, pos, cutoff, types=types)
    i, j, D = nl["i"], nl["j"], nl["D"]
    off = np.zeros(nat + 1, np.int64)
    np.add.at(off, i + 1, 1)
    off = np.cumsum(off)
    cnt[:] = np.diff(off)
    for a in range(nat):
        m = min(int(cnt[a]), K)
        idx[a, :m] = j[off[a]:off[a] + m]
        dist[a, :m] = D[off[a]:off[a] + m]
    return idx, dist, cnt, bool(np.any(cnt > K))


def ref_neighbour_list(quantities, cell_origin, cell, inv_cell, pbc, positions, cutoff,
                       types=None):
    """The compiled reference core (oracle/_ref), same argument handling."""
    lib = _load_ref()
    origin, cell, inv_cell, pbc, pos, ty = _prep(cell_origin, cell, inv_cell, pbc,
                                                 positions, types)
    nat = pos.shape[0]
    rc, per_atom, per_type_sq, nc = resolve_cutoff(cutoff, nat)
    flags = quantity_flags(quantities)
    res = _RefResult()
    st = lib.ref_neighbour_list(flags, _ptr(origin, "d"), _ptr(cell, "d"),
                                _ptr(inv_cell, "d"), _ptr(pbc, "i"), nat,
                                _ptr(pos, "d"), rc, _ptr(per_atom, "d"),
                                _ptr(per_type_sq, "d"), nc, _ptr(ty, "i"),
                                ctypes.byref(res))
    if st != 0:
        raise RuntimeError(lib.ref_last_error().decode())
    n = res.npairs
    out = {}
    try:
        if flags & 1:
            out["i"] = _copy(res.first, n, np.int32, (n,))
        if flags & 2:
            out["j"] = _copy(res.secnd, n, np.int32, (n,))
        if flags & 4:
            out["D"] = _copy(res.distvec, 3 * n, np.float64, (n, 3))
        if flags & 8:
            out["d"] = _copy(res.absdist, n, np.float64, (n,))
        if flags & 16:
            out["S"] = _copy(res.shift, 3 * n, np.int32, (n, 3))
    finally:
        lib.ref_free(ctypes.byref(res))
    return out


def ref_time(quantities, cell_origin, cell, inv_cell, pbc, positions, cutoff, n_rep,
             threads=None):
    """Wall seconds of n_rep calls of the compiled reference (scalar cutoff),
    timed inside the C++ library; returns (seconds[n_rep], npairs, threads)."""
    lib = _load_ref()
    if threads is not None:
        lib.ref_set_threads(int(threads))
    origin, cell, inv_cell, pbc, pos, _ = _prep(cell_origin, cell, inv_cell, pbc,
                                                positions, None)
    secs = np.zeros(n_rep, np.float64)
    npairs = ctypes.c_int64(0)
    st = lib.ref_time_neighbour_list(quantity_flags(quantities), _ptr(origin, "d"),
                                     _ptr(cell, "d"), _ptr(inv_cell, "d"),
                                     _ptr(pbc, "i"), pos.shape[0], _ptr(pos, "d"),
                                     float(cutoff), int(n_rep), _ptr(secs, "d"),
                                     ctypes.byref(npairs))
    if st != 0:
        raise RuntimeError(lib.ref_last_error().decode())
    return secs, int(npairs.value), int(lib.ref_max_threads())


def ref_first_neighbours(n, i):
    lib = _load_ref()
    i = np.ascontiguousarray(i, dtype=np.int32)
    seed = np.zeros(n + 1, np.int32)
    lib.ref_first_neighbours(int(n), int(i.shape[0]), _ptr(i, "i"), _ptr(seed, "i"))
    return seed


def ref_get_jump_indicies(sorted_arr):
    lib = _load_ref()
    s = np.ascontiguousarray(sorted_arr, dtype=np.int32)
    cap = s.shape[0] + 2
    out = np.zeros(cap, np.int32)
    n = lib.ref_get_jump_indicies(int(s.shape[0]), _ptr(s, "i"), _ptr(out, "i"), cap)
    return out[:n].copy()


def ref_triplet_list(first_i, absdist=None, cutoff=0.0):
    lib = _load_ref()
    f = np.ascontiguousarray(first_i, dtype=np.int32)
    d = None if absdist is None else np.ascontiguousarray(absdist, dtype=np.float64)
    n = lib.ref_triplet_list(int(f.shape[0]), _ptr(f, "i"), _ptr(d, "d"), float(cutoff),
                             None, None, 0)
    ij = np.zeros(max(n, 1), np.int32)
    ik = np.zeros(max(n, 1), np.int32)
    lib.ref_triplet_list(int(f.shape[0]), _ptr(f, "i"), _ptr(d, "d"), float(cutoff),
                         _ptr(ij, "i"), _ptr(ik, "i"), n)
    return ij[:n].copy(), ik[:n].copy()


def ref_neighbour_matrix(cell_origin, cell, inv_cell, pbc, positions, cutoff, K,
                         types=None):
    lib = _load_ref()
    origin, cell, inv_cell, pbc, pos, ty = _prep(cell_origin, cell, inv_cell, pbc,
                                                 positions, types)
    nat = pos.shape[0]
    rc, per_atom, per_type_sq, nc = resolve_cutoff(cutoff, nat)
    idx = np.zeros(nat * K, np.int32)
    dist = np.zeros(nat * K * 3, np.float64)
    cnt = np.zeros(nat, np.int32)
    ovf = np.zeros(1, np.int32)
    st = lib.ref_neighbour_matrix(_ptr(origin, "d"), _ptr(cell, "d"), _ptr(inv_cell, "d"),
                                  _ptr(pbc, "i"), nat, _ptr(pos, "d"), rc,
                                  _ptr(per_atom, "d"), _ptr(per_type_sq, "d"), nc,
                                  _ptr(ty, "i"), int(K), _ptr(idx, "i"), _ptr(dist, "d"),
                                  _ptr(cnt, "i"), _ptr(ovf, "i"))
    if st != 0:
        raise RuntimeError(lib.ref_last_error().decode())
    return idx.reshape(nat, K), dist.reshape(nat, K, 3), cnt, bool(ovf[0])
