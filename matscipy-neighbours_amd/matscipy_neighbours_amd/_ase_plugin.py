"""ASE neighbour-list plugin (the ``ase.plugins`` entry point).

Mirrors language_bindings/python/matscipy_neighbours/_ase_plugin.py of the
reference: the GPU neighbour list is made *available* to ASE under the name
"matscipy-neighbours-amd" (never selected automatically). With an ASE that
has no plugin API nothing is registered; the engine itself is imported only
when the adapter is called.
"""
from __future__ import annotations


def neighbor_list(quantities, atoms, cutoff, *, self_interaction=False):
    """The engine's neighbour list in ASE's calling convention (quantity
    letters i, j, d, D, S; flat arrays). Pure self pairs are never returned,
    so ``self_interaction=True`` is rejected rather than answered
    differently."""
    if self_interaction:
        raise NotImplementedError(
            "self_interaction=True is not available: matscipy-neighbours-amd never "
            "lists an atom as its own neighbour at zero shift")
    from .neighbours import neighbour_list as _neighbour_list

    return _neighbour_list(quantities, atoms, cutoff)


try:
    from ase._4.plugins.neighborlist import NeighborListPlugin
except ImportError:
    __ase_plugins__: set = set()
else:  # pragma: no cover - needs an ASE with the plugin API
    __ase_plugins__ = {
        NeighborListPlugin(
            "matscipy-neighbours-amd",
            long_name="matscipy-neighbours on AMD Instinct GPUs (HIP)",
            citation="https://github.com/libAtoms/matscipy-neighbours",
            implementation="matscipy_neighbours_amd._ase_plugin.neighbor_list",
        ),
    }
