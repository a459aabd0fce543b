"""Host-side argument handling of the Python API, on the CPU: device routing
(device= and where the positions live), the geometry defaults and the
shrink-wrapped box, cutoff dicts, the ASE-or-arrays gathering, mic and the
DLPack wrapper. Each expectation is the behaviour of the reference's
language_bindings/python/matscipy_neighbours/neighbours.py:39-197, written
out as values."""

import numpy as np
import pytest

from matscipy_neighbours_amd import DLPackTensor, mic
from matscipy_neighbours_amd import neighbours as nb


class _FakeArray:
    def __init__(self, dev):
        self._dev = dev

    def __dlpack_device__(self):
        return self._dev


class _Broken:
    def __dlpack_device__(self):
        raise RuntimeError("no device")


def test_dlpack_device_and_location():
    assert nb._dlpack_device(np.zeros(3)) == (1, 0)
    assert nb._dlpack_device([1, 2, 3]) is None
    assert nb._dlpack_device(_Broken()) is None
    assert nb._dlpack_device(_FakeArray((10, 3))) == (10, 3)
    assert not nb._is_on_device(np.zeros(3))
    assert not nb._is_on_device([0.0])
    assert not nb._is_on_device(_Broken())
    assert nb._is_on_device(_FakeArray((10, 0)))   # ROCm
    assert nb._is_on_device(_FakeArray((2, 1)))    # CUDA
    assert not nb._is_on_device(_FakeArray((1, 0)))


@pytest.mark.parametrize("device,on_dev,want", [
    (None, False, (False, -1)), (None, True, (True, -1)),
    ("cpu", True, (False, -1)), ("CPU", False, (False, -1)),
    ("cuda", False, (True, -1)), ("GPU", False, (True, -1)), ("rocm", False, (True, -1)),
    ("hip", True, (True, -1)), (2, False, (True, 2)), (("cuda", 1), False, (True, 1)),
    (["rocm", 3], True, (True, 3)),
])
def test_device_routing(device, on_dev, want):
    assert nb._resolve_device(device, on_dev) == want


@pytest.mark.parametrize("bad", ["tpu", 1.5, ("cuda",), ("a", 1, 2), object()])
def test_device_routing_rejects(bad):
    with pytest.raises(ValueError, match="Unknown device"):
        nb._resolve_device(bad, False)


def test_geometry_defaults_and_shrink_wrap():
    pos = np.array([[1.0, -2.0, 0.5], [3.0, 1.0, 4.5], [2.0, 0.0, 1.0]])
    co, ce, inv, pbc, nums = nb._host_metadata(None, None, None, None, 3, positions=pos)
    assert np.array_equal(co, [1.0, -2.0, 0.5])          # minimum corner
    assert np.array_equal(ce, np.diag([2.0, 3.0, 4.0]))  # extent of the atoms
    assert np.array_equal(pbc, [False, False, False])
    assert nums.dtype == np.int32 and np.array_equal(nums, [1, 1, 1])
    assert np.allclose(inv @ ce.T, np.eye(3))
    # a given origin is kept with the shrink-wrapped box
    co2, ce2, *_ = nb._host_metadata(None, None, None, [0.0, 0.0, 0.0], 3, positions=pos)
    assert np.array_equal(co2, [0.0, 0.0, 0.0]) and np.array_equal(ce2, ce)
    # explicit geometry: origin 0, pbc broadcast, inv = inv(cell.T), all contiguous
    cell = np.array([[4.0, 0.0, 0.0], [1.0, 5.0, 0.0], [0.5, 0.2, 6.0]])
    co, ce, inv, pbc, nums = nb._host_metadata(cell.T.T, True, [3, 1], None, 2)
    assert np.array_equal(co, np.zeros(3)) and np.array_equal(pbc, [True, True, True])
    assert np.array_equal(inv, np.linalg.inv(cell.T))
    assert np.array_equal(nums, [3, 1]) and nums.dtype == np.int32
    for a in (co, ce, inv, pbc, nums):
        assert a.flags.c_contiguous
    with pytest.raises(ValueError):
        nb._host_metadata(None, None, None, None, 3)


def test_cutoff_dict_to_matrix():
    nums = np.array([1, 8, 8, 1], np.int32)
    m, n2 = nb._resolve_cutoff({("H", "O"): 1.2, (8, 8): 2.5, ("H", "Fe"): 9.0}, nums)
    assert n2 is nums and m.shape == (9, 9)
    assert m[1, 8] == m[8, 1] == 1.2 and m[8, 8] == 2.5
    assert np.count_nonzero(m) == 3                  # Fe (26) > max number: dropped
    assert nb._resolve_cutoff(3.0, nums) == (3.0, nums)


def test_gather_forms():
    pos = np.zeros((2, 3))
    assert nb._gather(None, pos, None, None, None, None)[0] is pos
    with pytest.raises(ValueError, match="Provide either"):
        nb._gather(None, None, None, None, None, None)

    class Atoms:
        positions = pos
        cell = np.eye(3)
        pbc = np.array([True, False, True])
        numbers = np.array([1, 6])

    p, c, b, n, o = nb._gather(Atoms(), None, None, None, None, None)
    assert p is pos and np.array_equal(c, np.eye(3)) and n.dtype == np.int32
    assert np.array_equal(o, np.zeros(3))
    with pytest.raises(ValueError, match="Cannot combine"):
        nb._gather(Atoms(), pos, None, None, None, None)


def test_mic_values():
    cell = np.array([[10.0, 0.0, 0.0], [0.0, 8.0, 0.0], [0.0, 0.0, 6.0]])
    dr = np.array([[6.0, -5.0, 3.0], [-9.5, 3.9, -3.1], [5.0, 4.0, 3.0]])
    # round half to even at exactly half a period
    assert np.array_equal(mic(dr, cell), [[-4.0, 3.0, 3.0], [0.5, 3.9, 2.9], [5.0, 4.0, 3.0]])
    # pbc masks the Cartesian components entering the fractional transform
    assert np.array_equal(mic(dr, cell, [False, True, True]),
                          [[6.0, 3.0, 3.0], [-9.5, 3.9, 2.9], [5.0, 4.0, 3.0]])
    # the mask scales rows of inv(cell), i.e. Cartesian components of dr (a
    # column mask, on fractional coordinates, would give [10, 3, 0] here)
    skew = np.array([[5.0, 1.0, 0.0], [0.0, 5.0, 0.0], [0.0, 0.0, 5.0]])
    assert np.allclose(mic([[10.0, 3.0, 0.0]], skew, [False, True, True]), [[10.0, -2.0, 0.0]])
    tri = np.array([[5.0, 0.0, 0.0], [2.0, 5.0, 0.0], [0.0, 0.0, 5.0]])
    out = mic([[6.0, 5.5, 0.0]], tri)
    assert np.allclose(out, [[-1.0, 0.5, 0.0]])


def test_dlpack_wrapper_is_the_capsule():
    cap = object()
    t = DLPackTensor(cap, (10, 0))
    assert t.__dlpack__() is cap and t.__dlpack__(stream=None) is cap
    assert t.__dlpack_device__() == (10, 0)
