"""The C ABI on the GPU: one-shot entry points, host copies, DLPack export,
count-only, error reporting, and the device square root."""

import ctypes

import numpy as np
import pytest

import oracle
from conftest import assert_same_pairs, bits

pytestmark = pytest.mark.gpu


def _lib():
    from matscipy_neighbours_amd import _lib as L
    return L.load()


def _geom(cell, pbc):
    cell = np.ascontiguousarray(cell, dtype=np.float64)
    inv = np.ascontiguousarray(np.linalg.inv(cell.T))
    pb = np.ascontiguousarray(np.broadcast_to(np.asarray(pbc, bool), (3,)), dtype=np.int32)
    return np.zeros(3), cell, inv, pb


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def test_device_sqrt_matches_host(gpu):
    """d is sqrt(r2) on the device; it must round exactly like the host's
    IEEE sqrt. Random magnitudes, squares of representable values, and
    subnormal/huge edge values."""
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.uniform(0, 50, 200000),
        np.exp(rng.uniform(-700, 700, 100000)),
        rng.uniform(0, 1, 1000) ** 2,
        np.array([0.0, 1.0, 2.0, 4.0, 1e-310, 5e-324, 1.7e308, 25.0, 24.999999999999996]),
        np.nextafter(np.arange(1, 1000, dtype=float) ** 2, 0),
    ])
    y = np.empty_like(x)
    assert _lib().mnl_debug_sqrt(_p(x), _p(y), x.size) == 0
    assert np.array_equal(bits(y), bits(np.sqrt(x)))


def test_one_shot_host_in_host_out(gpu):
    L = _lib()
    rng = np.random.default_rng(3)
    pos = np.ascontiguousarray(rng.uniform(0, 12, (2000, 3)))
    co, cell, inv, pb = _geom(12 * np.eye(3), True)
    res = ctypes.c_void_p()
    st = L.mnl_neighbour_list_device(31, _p(co), _p(cell), _p(inv), _p(pb), 2000, _p(pos), 0,
                                     1.0, None, None, 0, None, -1, ctypes.byref(res))
    assert st == 0, L.mnl_last_error()
    try:
        P = L.mnl_result_npairs(res)
        ref = oracle.neighbour_list("ijdDS", co, cell, inv, pb, pos, 1.0)
        assert P == len(ref["i"])
        got = {"i": np.empty(P, np.int32), "j": np.empty(P, np.int32),
               "d": np.empty(P), "D": np.empty((P, 3)), "S": np.empty((P, 3), np.int32)}
        for q, a in got.items():
            assert L.mnl_result_copy_to_host(res, q.encode(), _p(a)) == 0
        assert_same_pairs(got, ref)
        assert L.mnl_result_nat(res) == 2000
    finally:
        L.mnl_result_free(res)


def test_dlpack_export_zero_copy(gpu):
    """Export through DLPack and adopt in torch without a copy."""
    import torch
    L = _lib()
    rng = np.random.default_rng(4)
    pos = np.ascontiguousarray(rng.uniform(0, 10, (1000, 3)))
    co, cell, inv, pb = _geom(10 * np.eye(3), True)
    res = ctypes.c_void_p()
    assert L.mnl_neighbour_list_device(31, _p(co), _p(cell), _p(inv), _p(pb), 1000, _p(pos),
                                       0, 2.5, None, None, 0, None, -1,
                                       ctypes.byref(res)) == 0
    ref = oracle.neighbour_list("ijdDS", co, cell, inv, pb, pos, 2.5)
    PyCapsule_New = ctypes.pythonapi.PyCapsule_New
    PyCapsule_New.restype = ctypes.py_object
    PyCapsule_New.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
    try:
        for q in "ijDdS":
            dev_ptr = L.mnl_result_data(res, q.encode())
            t_ptr = L.mnl_result_export_dlpack(res, q.encode())
            assert t_ptr, L.mnl_last_error()
            assert L.mnl_result_data(res, q.encode()) is None
            cap = PyCapsule_New(t_ptr, b"dltensor", None)
            t = torch.from_dlpack(cap)
            assert t.is_cuda and t.data_ptr() == dev_ptr
            assert np.array_equal(bits(t.cpu().numpy()), bits(ref[q]))
            del t
        assert not L.mnl_result_export_dlpack(res, b"i")
    finally:
        L.mnl_result_free(res)


def test_count_only(gpu):
    L = _lib()
    rng = np.random.default_rng(5)
    pos = np.ascontiguousarray(rng.uniform(0, 14, (3000, 3)))
    co, cell, inv, pb = _geom(14 * np.eye(3), True)
    res = ctypes.c_void_p()
    assert L.mnl_neighbour_count_device(_p(co), _p(cell), _p(inv), _p(pb), 3000, _p(pos), 0,
                                        1.2, None, None, 0, None, -1,
                                        ctypes.byref(res)) == 0
    try:
        c = np.empty(3000, np.int32)
        assert L.mnl_result_copy_to_host(res, b"c", _p(c)) == 0
        ref = oracle.neighbour_list("i", co, cell, inv, pb, pos, 1.2, with_counts=True)
        assert np.array_equal(c, ref["counts"])
        assert L.mnl_result_npairs(res) == len(ref["i"])
        assert L.mnl_result_data(res, b"i") is None
    finally:
        L.mnl_result_free(res)


def test_zero_volume_error(gpu):
    """tests/test_core.cc:64-73: NL_ERROR and a message, not a crash."""
    L = _lib()
    bad = np.ascontiguousarray([[1.0, 0, 0], [2.0, 0, 0], [0, 0, 1.0]])
    inv = np.ascontiguousarray(np.eye(3))
    pb = np.ones(3, np.int32)
    pos = np.array([[0.5, 0.5, 0.5]])
    res = ctypes.c_void_p()
    st = L.mnl_neighbour_list_device(1, _p(np.zeros(3)), _p(bad), _p(inv), _p(pb), 1,
                                     _p(pos), 0, 1.0, None, None, 0, None, -1,
                                     ctypes.byref(res))
    assert st == -1
    assert L.mnl_last_error() == b"Zero cell volume."
    assert not res.value


def _block_bytes(b):
    """The library's size classes for one-shot buffers: b rounded up to 1/8
    of its highest power of two, at least 64 KiB."""
    top = 1
    while top <= b // 2:
        top *= 2
    g = max(64 << 10, top // 8)
    return (b + g - 1) // g * g


def test_result_buffers_reused(gpu):
    """Freed one-shot outputs (by mnl_result_free or by an exported tensor's
    DLPack deleter) serve the next result whose buffers fall in the same size
    classes: the same device buffers come back, for the same input and for a
    jittered copy with 46 more pairs (an MD step), and every result still
    matches the oracle bit for bit (a kept buffer's old contents never show
    through). After mnl_release_cached_memory the builds allocate afresh and
    stay exact."""
    import torch
    L = _lib()
    L.mnl_release_cached_memory()
    rng = np.random.default_rng(6)
    pos_a = np.ascontiguousarray(rng.uniform(0, 11, (1500, 3)))
    pos_b = np.ascontiguousarray(pos_a + np.random.default_rng(7).normal(0, 0.02, pos_a.shape))
    co, cell, inv, pb = _geom(11 * np.eye(3), True)
    refs = [oracle.neighbour_list("ijdDS", co, cell, inv, pb, p, 2.0) for p in (pos_a, pos_b)]
    Pa, Pb = (len(r["i"]) for r in refs)
    assert Pa != Pb and all(_block_bytes(Pa * k) == _block_bytes(Pb * k) for k in (4, 8, 12, 24))

    def build(which):
        pos, ref = (pos_a, pos_b)[which], refs[which]
        P = len(ref["i"])
        res = ctypes.c_void_p()
        assert L.mnl_neighbour_list_device(31, _p(co), _p(cell), _p(inv), _p(pb), 1500,
                                           _p(pos), 0, 2.0, None, None, 0, None, -1,
                                           ctypes.byref(res)) == 0, L.mnl_last_error()
        got = {"i": np.empty(P, np.int32), "j": np.empty(P, np.int32),
               "d": np.empty(P), "D": np.empty((P, 3)), "S": np.empty((P, 3), np.int32)}
        assert L.mnl_result_npairs(res) == P
        for q, a in got.items():
            assert L.mnl_result_copy_to_host(res, q.encode(), _p(a)) == 0
        assert_same_pairs(got, ref)
        ptrs = {q: L.mnl_result_data(res, q.encode()) for q in "ijDdSc"}
        return res, ptrs

    res, first = build(0)
    L.mnl_result_free(res)
    res, second = build(0)
    # i and j have the same size and may trade buffers
    assert sorted(first.values()) == sorted(second.values())
    assert second["D"] == first["D"] and second["S"] == first["S"]
    L.mnl_result_free(res)
    res, jittered = build(1)
    assert sorted(jittered.values()) == sorted(first.values())

    PyCapsule_New = ctypes.pythonapi.PyCapsule_New
    PyCapsule_New.restype = ctypes.py_object
    PyCapsule_New.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
    t = torch.from_dlpack(PyCapsule_New(L.mnl_result_export_dlpack(res, b"D"), b"dltensor",
                                        None))
    assert t.data_ptr() == jittered["D"]
    L.mnl_result_free(res)
    t.add_(1.0)  # the exported buffer outlives its result
    del t
    torch.cuda.synchronize()
    res, third = build(0)
    assert third["D"] == jittered["D"]
    L.mnl_result_free(res)

    L.mnl_release_cached_memory()
    res, _ = build(1)
    L.mnl_result_free(res)
    L.mnl_release_cached_memory()


def test_kept_buffers_released_when_device_full(gpu):
    """A one-shot build that finds the device out of memory first returns
    the buffers the library keeps, then allocates again. 200,000 atoms of a
    uniform gas (0.1 atoms/A^3): a build at rc 5 is freed (its ~0.55 GB of
    outputs kept). torch then takes the device's memory in 64 MiB pieces
    until it runs out and gives back 9 of them (0.6 GB), less than the build
    at rc 6 needs (~0.98 GB at its size classes), so its outputs fit only
    once the kept ones are gone, which the cache's byte count shows. The rc 6
    list is checked on the device against D == r[j] - r[i] + S L, d == |D| <
    rc, and its length against the count pass."""
    import torch
    L = _lib()
    L.mnl_release_cached_memory()
    n = 200_000
    box = (n / 0.1) ** (1.0 / 3.0)
    rng = np.random.default_rng(8)
    r = torch.from_numpy(rng.uniform(0.0, box, (n, 3))).to("cuda:0")
    co, cell, inv, pb = _geom(box * np.eye(3), True)

    def build(rc):
        res = ctypes.c_void_p()
        st = L.mnl_neighbour_list_device(31, _p(co), _p(cell), _p(inv), _p(pb), n,
                                         ctypes.c_void_p(r.data_ptr()), 1, rc, None, None, 0,
                                         None, 0, ctypes.byref(res))
        return st, res

    st, res = build(5.0)
    assert st == 0, L.mnl_last_error()
    L.mnl_result_free(res)
    cres = ctypes.c_void_p()
    assert L.mnl_neighbour_count_device(_p(co), _p(cell), _p(inv), _p(pb), n,
                                        ctypes.c_void_p(r.data_ptr()), 1, 6.0, None, None, 0,
                                        None, 0, ctypes.byref(cres)) == 0
    P6 = L.mnl_result_npairs(cres)
    L.mnl_result_free(cres)
    kept = L.mnl_cached_bytes(0)
    assert kept > 0.5e9
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    free, _ = torch.cuda.mem_get_info()
    piece = 64 << 20
    filler = [torch.empty(free - (2 << 30), dtype=torch.uint8, device="cuda:0")]
    try:
        while True:
            try:
                filler.append(torch.empty(piece, dtype=torch.uint8, device="cuda:0"))
            except torch.OutOfMemoryError:
                break
        del filler[-9:]
        torch.cuda.empty_cache()
        print(f"{P6} pairs at rc 6, {kept / 2**30:.2f} GiB kept, "
              f"{len(filler) - 1} pieces held", flush=True)
        st, res = build(6.0)
        assert st == 0, L.mnl_last_error()
        assert L.mnl_cached_bytes(0) == 0
    finally:
        del filler
        torch.cuda.empty_cache()
    try:
        assert L.mnl_result_npairs(res) == P6
        ptr = {q: L.mnl_result_data(res, q.encode()) for q in "ijDdS"}
        i, j, D, d, S = (torch.from_dlpack(_capsule(L, res, q)) for q in "ijDdS")
        assert i.data_ptr() == ptr["i"] and D.data_ptr() == ptr["D"]
        ref_D = r[j.long()] - r[i.long()] + S.double() * box
        assert float((D - ref_D).abs().max()) < 1e-9
        assert float((d - torch.linalg.norm(D, dim=1)).abs().max()) < 1e-12
        assert bool((d < 6.0).all()) and bool((i[1:] >= i[:-1]).all())
        del i, j, D, d, S, ref_D
    finally:
        L.mnl_result_free(res)
    torch.cuda.synchronize()
    L.mnl_release_cached_memory()


def _capsule(L, res, q):
    PyCapsule_New = ctypes.pythonapi.PyCapsule_New
    PyCapsule_New.restype = ctypes.py_object
    PyCapsule_New.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
    t = L.mnl_result_export_dlpack(res, q.encode())
    assert t, L.mnl_last_error()
    return PyCapsule_New(t, b"dltensor", None)
