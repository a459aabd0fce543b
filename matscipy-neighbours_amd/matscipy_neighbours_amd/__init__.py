"""matscipy_neighbours_amd - MI355X-native neighbour lists.

A drop-in for the public API of ``matscipy_neighbours`` (itself compatible
with ``matscipy.neighbours``) whose every build runs in hand-written gfx950
HIP kernels (libmnl.so, C ABI in include/mnl.h). Pairs, their order and the
bits of D and d are identical to the reference CPU path.
"""

from .autograd import neighbour_list_autograd
from .neighbours import (
    DLPackTensor,
    autotune,
    autotune_batch,
    coordination,
    first_neighbours,
    get_jump_indicies,
    lennard_jones,
    mic,
    neighbour_list,
    neighbour_list_batch,
    neighbour_matrix,
    triplet_list,
)

__all__ = [
    "neighbour_list",
    "neighbour_list_batch",
    "neighbour_list_autograd",
    "neighbour_matrix",
    "first_neighbours",
    "get_jump_indicies",
    "triplet_list",
    "mic",
    "coordination",
    "lennard_jones",
    "autotune",
    "autotune_batch",
    "DLPackTensor",
]
