"""Host-side API checks that need no GPU: exports, the ASE plugin module,
argument errors raised before any device work (neighbours.py and
bind_py_neighbours.cc of the reference)."""

import numpy as np
import pytest


def test_public_names_match_reference():
    import matscipy_neighbours_amd as m
    for name in ("neighbour_list", "neighbour_matrix", "first_neighbours", "get_jump_indicies",
                 "triplet_list", "mic", "coordination", "DLPackTensor"):
        assert name in m.__all__ and hasattr(m, name)


def test_ase_plugin_module():
    from matscipy_neighbours_amd import _ase_plugin
    assert isinstance(_ase_plugin.__ase_plugins__, set)
    with pytest.raises(NotImplementedError):
        _ase_plugin.neighbor_list("ij", None, 1.0, self_interaction=True)


def test_mic():
    from matscipy_neighbours_amd import mic
    cell = np.diag([10.0, 10.0, 10.0])
    dr = np.array([[9.0, -6.0, 0.5]])
    assert np.allclose(mic(dr, cell), [[-1.0, 4.0, 0.5]])
    assert np.allclose(mic(dr, cell, pbc=[False, True, True]), [[9.0, 4.0, 0.5]])


def test_argument_errors_before_device_work():
    from matscipy_neighbours_amd import neighbour_list
    with pytest.raises(ValueError, match="cutoff"):
        neighbour_list("ij", positions=np.zeros((2, 3)), cell=np.eye(3))
    with pytest.raises(ValueError, match="Unsupported quantity"):
        neighbour_list("ijx", positions=np.zeros((2, 3)), cell=np.eye(3), cutoff=1.0)


def test_lennard_jones_argument_errors():
    from matscipy_neighbours_amd import lennard_jones
    with pytest.raises(ValueError, match="cutoff"):
        lennard_jones(positions=np.zeros((2, 3)), cell=np.eye(3))
    with pytest.raises(TypeError, match="single float cutoff"):
        lennard_jones(positions=np.zeros((2, 3)), cell=np.eye(3), cutoff=np.array([1.0, 1.0]))
    with pytest.raises(TypeError, match="single float cutoff"):
        lennard_jones(positions=np.zeros((2, 3)), cell=np.eye(3), cutoff={(1, 1): 1.0})
