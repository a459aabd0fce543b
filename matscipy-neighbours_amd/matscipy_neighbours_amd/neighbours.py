"""Public neighbour-list API, a drop-in for ``matscipy_neighbours``.

Mirrors language_bindings/python/matscipy_neighbours/neighbours.py of the
reference: same functions, argument meaning, return layout and error types.
The computation always runs in the mnl HIP engine (libmnl.so) on the GPU;
PyTorch only provides device memory and the current stream. There is no CPU
code path: without a HIP device every call raises RuntimeError.

Differences from the reference, all deliberate:

* host (numpy) inputs are computed on the GPU too and come back as numpy
  arrays, like the reference's host fast path; ``device="cpu"`` asks for
  host-resident outputs, it does not select a CPU implementation;
* device outputs are torch tensors by default (the reference returns cupy,
  which is absent from ROCm PyTorch images);
* ``out_dtype=numpy.float32`` (or torch.float32) writes D and d in float32;
  the pair set is still decided in float64.
"""

import threading

import numpy as np

from . import _lib

__all__ = [
    "neighbour_list",
    "neighbour_list_batch",
    "neighbour_matrix",
    "coordination",
    "first_neighbours",
    "get_jump_indicies",
    "triplet_list",
    "mic",
    "lennard_jones",
    "autotune",
    "autotune_batch",
    "DLPackTensor",
]

_DLPACK_CPU = 1
_DLPACK_CUDA = 2
_DLPACK_ROCM = 10

# Chemical symbols by atomic number, for element-pair cutoff dicts
# (the reference uses ase.data.atomic_numbers when ASE is installed).
_SYMBOLS = (
    "X H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co "
    "Ni Cu Zn Ga Ge As Se Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te "
    "I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re Os Ir "
    "Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No "
    "Lr Rf Db Sg Bh Hs Mt Ds Rg Cn Nh Fl Mc Lv Ts Og").split()
try:  # pragma: no cover - ase is optional
    from ase.data import atomic_numbers as _atomic_numbers
except ImportError:
    _atomic_numbers = {s: z for z, s in enumerate(_SYMBOLS)}


def _torch():
    import torch
    return torch


def mic(dr, cell, pbc=None):
    """Minimum-image form of distance vectors ``dr`` (n x 3) in ``cell``
    (lattice vectors as rows), as language_bindings/python/
    matscipy_neighbours/neighbours.py:39-62 defines it: the image count is
    the rounded (half to even) fractional coordinate of ``dr``, and ``pbc``
    (default: every direction periodic) zeroes the Cartesian components of
    ``dr`` that enter that fractional transform, exactly as the reference
    scales the rows of the inverse cell."""
    vectors = np.asarray(dr, dtype=np.float64)
    lattice = np.asarray(cell, dtype=np.float64)
    to_frac = np.linalg.inv(lattice)
    if pbc is not None:
        to_frac = to_frac * np.asarray(pbc, dtype=np.int64).reshape(3, 1)
    images = np.rint(vectors @ to_frac)
    return vectors - images @ lattice


class DLPackTensor:
    """One exported quantity (reference neighbours.py:65-81): an owning
    "dltensor" capsule plus its (device type, device id). Any DLPack consumer
    adopts it with ``from_dlpack``. The capsule can be consumed once, so the
    object is single use, as in the reference."""

    __slots__ = ("_capsule", "_device")

    def __init__(self, capsule, device):
        self._capsule, self._device = capsule, tuple(device)

    def __dlpack__(self, stream=None, **_ignored):
        return self._capsule

    def __dlpack_device__(self):
        return self._device


def _dlpack_device(x):
    """(device type, device id) that ``x`` reports through DLPack; None when
    it has no ``__dlpack_device__`` or the call fails."""
    query = getattr(x, "__dlpack_device__", None)
    if not callable(query):
        return None
    try:
        kind, index = query()
    except Exception:
        return None
    return int(kind), int(index)


def _is_on_device(x):
    """Whether ``x`` is an array in GPU memory (any DLPack device but the host)."""
    where = _dlpack_device(x)
    return where is not None and where[0] != _DLPACK_CPU


# device= spellings that select the GPU (reference neighbours.py:111-128)
_GPU_NAMES = frozenset(("cuda", "gpu", "rocm", "hip"))


def _resolve_device(device, on_device_input):
    """(device outputs wanted, device id; -1 = current) for the ``device=``
    argument: None follows where the positions are; "cpu" asks for host
    outputs; a GPU name, an int id or a (name, id) pair asks for device
    outputs. Anything else is a ValueError."""
    if device is None:
        return bool(on_device_input), -1
    if isinstance(device, str):
        name = device.lower()
        if name == "cpu":
            return False, -1
        if name in _GPU_NAMES:
            return True, -1
    elif isinstance(device, int):
        return True, device
    elif isinstance(device, (tuple, list)) and len(device) == 2:
        return True, int(device[1])
    raise ValueError(f"Unknown device {device!r}.")


def _host_metadata(cell, pbc, numbers, cell_origin, nat, *, positions=None):
    """The host geometry of a build (reference neighbours.py:145-164):
    origin 0, pbc all False and numbers all 1 by default; without a cell the
    box shrink-wraps the (host) positions, origin at their minimum corner
    unless given; inv_cell = inv(cell.T). Returns contiguous (origin, cell,
    inv_cell, pbc bool[3], numbers int32)."""
    if cell is None:
        if positions is None:
            raise ValueError("cell must be given for device-resident positions.")
        pts = np.asarray(positions, dtype=np.float64)
        lo = pts.min(axis=0)
        cell = np.diag(pts.max(axis=0) - lo)
        if cell_origin is None:
            cell_origin = lo
    origin = np.zeros(3) if cell_origin is None else cell_origin
    periodic = np.zeros(3, dtype=bool) if pbc is None else pbc
    types = np.ones(nat, dtype=np.int32) if numbers is None else numbers
    cell = np.ascontiguousarray(cell, dtype=np.float64)
    return (np.ascontiguousarray(origin, dtype=np.float64),
            cell,
            np.ascontiguousarray(np.linalg.inv(cell.T)),
            np.ascontiguousarray(np.broadcast_to(periodic, (3,)), dtype=bool),
            np.ascontiguousarray(types, dtype=np.int32))


def _element_number(key):
    """Atomic number of a cutoff-dict key: a chemical symbol or a number."""
    return _atomic_numbers.get(key, key)


def _resolve_cutoff(cutoff, numbers):
    """A cutoff dict keyed by element pairs becomes the symmetric per-type
    matrix of side max(numbers) + 1 (pairs naming a larger element are
    dropped), as in reference neighbours.py:167-181; any other cutoff passes
    through. Returns (cutoff, numbers)."""
    if not isinstance(cutoff, dict):
        return cutoff, numbers
    side = int(np.max(numbers)) + 1
    table = np.zeros((side, side))
    for pair, value in cutoff.items():
        za, zb = (_element_number(e) for e in pair)
        if za < side and zb < side:
            table[za, zb] = table[zb, za] = value
    return table, numbers


def _parse_cutoff(cutoff, nat):
    """The C binding's cutoff forms (bind_py_neighbours.cc:130-173): a Python
    float is the global cutoff; a 1-d array holds per-atom radii (binning
    cutoff = max 2r, length must be nat: TypeError); a 2-d square array is the
    per-type matrix (binning cutoff = max, squares used; non-square:
    TypeError). Like the reference, a Python int is not a scalar: it fails the
    array conversion with ValueError."""
    if isinstance(cutoff, float):
        return float(cutoff), None, None, 0
    a = np.asarray(cutoff, dtype=np.float64)
    if a.ndim not in (1, 2):
        raise ValueError("cutoff must be a float, a 1-d per-atom array or a 2-d "
                         "per-type array")
    a = np.ascontiguousarray(a)
    if a.ndim == 1:
        if a.shape[0] != nat:
            raise TypeError("One-dimensional cutoff array must have length that "
                            "corresponds to position array.")
        rc = 0.0
        for v in a:
            rc = max(rc, 2 * float(v))
        return rc, a, None, 0
    if a.shape[0] != a.shape[1]:
        raise TypeError("Two-dimensional cutoff array must be square.")
    rc = 0.0
    for v in a.ravel():
        rc = max(rc, float(v))
    return rc, None, np.ascontiguousarray(a.ravel() * a.ravel()), int(a.shape[0])


def _quantity_flags(quantities):
    if not isinstance(quantities, str):
        raise TypeError("quantities must be a str")
    flags = 0
    for q in quantities:
        if q not in _lib.QUANTITY_BITS:
            raise ValueError("Unsupported quantity specified.")
        flags |= _lib.QUANTITY_BITS[q]
    return flags


def _gather(atoms, positions, cell, pbc, numbers, cell_origin):
    """(positions, cell, pbc, numbers, cell_origin) from an ASE ``Atoms`` or
    from the explicit arguments (reference neighbours.py:184-197); device
    positions pass through untouched. Mixing the two forms, or giving
    neither, is a ValueError."""
    explicit = (positions, cell, pbc, numbers, cell_origin)
    if atoms is None:
        if positions is None:
            raise ValueError("Provide either an ASE Atoms object or a positions array.")
        return explicit
    if any(v is not None for v in explicit):
        raise ValueError("Cannot combine an ASE Atoms object with explicit "
                         "positions/cell/pbc/numbers/cell_origin.")
    return (atoms.positions, np.asarray(atoms.cell), atoms.pbc,
            np.asarray(atoms.numbers).astype(np.int32), np.zeros(3))


def _device_positions(positions):
    """A device array as a torch float64 (n, 3) contiguous tensor (zero copy
    through DLPack when it is not a torch tensor already)."""
    torch = _torch()
    t = positions if isinstance(positions, torch.Tensor) else torch.from_dlpack(positions)
    if t.dtype != torch.float64 or t.dim() != 2 or t.shape[1] != 3 or not t.is_contiguous():
        raise TypeError("device positions must be a C-contiguous float64 array of "
                        "shape (n, 3)")
    return t


class _Build:
    """One engine build on a device: count (binning + count + scan) then the
    requested fills, with outputs allocated from torch's caching allocator."""

    def __init__(self, positions_t, co, ce, inv, pbc, cutoff, numbers, backend=-1):
        torch = _torch()
        self.torch = torch
        self.dev = positions_t.device
        self.nat = int(positions_t.shape[0])
        rc, per_atom, per_type_sq, nc = _parse_cutoff(cutoff, self.nat)
        self.rc = rc
        self.per_atom_t = (None if per_atom is None else
                           torch.as_tensor(per_atom, dtype=torch.float64, device=self.dev))
        self.per_type_sq = per_type_sq
        self.nc = nc
        self.types_t = None
        if per_type_sq is not None:
            self.types_t = torch.as_tensor(np.asarray(numbers, dtype=np.int32),
                                           dtype=torch.int32, device=self.dev)
            if self.types_t.shape[0] != self.nat:
                raise TypeError("Position and type arrays must have identical first "
                                "dimension.")
        self.co = np.ascontiguousarray(co, dtype=np.float64)
        self.ce = np.ascontiguousarray(ce, dtype=np.float64)
        self.inv = np.ascontiguousarray(inv, dtype=np.float64)
        self.pbc = np.ascontiguousarray(np.asarray(pbc, dtype=bool), dtype=np.int32)
        self.pos = positions_t
        self.backend = backend

    def run(self, flags, want_counts=False, f32=False, packed=False, need_total=True):
        """Count, then fill the quantities in ``flags``. With ``packed`` the
        outputs are views of one device buffer (``out["_packed"]``: D and d
        first, then i, j, S, each part aligned to its element size), so that a
        host result needs a single transfer. ``need_total=False`` (no
        quantities) leaves the pair total unread, so the call does not wait
        for the device.

        Device outputs of a build like the plan's previous one (same atom
        count and cutoff, as in an MD loop): the fill is queued right behind
        the count into buffers sized from the previous pair total (plus 1%
        and 4096), and the total is read once, after the fill; the outputs
        are the first P rows of those buffers. A total above that capacity
        fills again into buffers of the exact size. This saves the wait
        between count and fill. Host results (``packed``) keep the exact
        sizes, since their copy to the host dominates."""
        torch = self.torch
        lib = _lib.load()
        idx = self.dev.index if self.dev.index is not None else torch.cuda.current_device()
        plan = _lib.plan_for(idx)
        out = {}
        with plan.lock, torch.cuda.device(idx):
            stream = torch.cuda.current_stream(idx).cuda_stream
            sp = _lib.ctypes.c_void_p(stream)
            args = (plan.handle, sp, _lib.ptr(self.co),
                    _lib.ptr(self.ce), _lib.ptr(self.inv), _lib.ptr(self.pbc), self.nat,
                    _lib.ptr(self.pos), self.rc, _lib.ptr(self.per_atom_t),
                    _lib.ptr(self.per_type_sq), self.nc, _lib.ptr(self.types_t),
                    int(self.backend))
            fdt = torch.float32 if f32 else torch.float64
            kw = dict(device=self.dev)
            key = (self.nat, repr(self.rc))
            last = getattr(plan, "last_build", None)
            if flags and not packed and last is not None and last[0] == key:
                np_dev = torch.empty(1, dtype=torch.int64, **kw)
                _lib.check(lib.mnl_plan_count_async(*args, _lib.ptr(np_dev)))
                cap = last[1] + last[1] // 100 + 4096
                bufs = self._alloc(flags, cap, fdt, kw, packed=False)
                self._fill_async(lib, plan, sp, cap, bufs, f32)
                P = int(np_dev.item())
                if P > cap:
                    bufs = self._alloc(flags, P, fdt, kw, packed=False)
                    self._fill_async(lib, plan, sp, P, bufs, f32)
                for q, t in bufs.items():
                    out[q] = t[:P]
            elif flags or need_total:
                npairs = _lib.ctypes.c_int64(0)
                _lib.check(lib.mnl_plan_count(*args, _lib.ctypes.byref(npairs)))
                P = int(npairs.value)
                out = self._alloc(flags, P, fdt, kw, packed)
                if flags:
                    fill = lib.mnl_plan_fill_f32 if f32 else lib.mnl_plan_fill
                    _lib.check(fill(plan.handle, sp, _lib.ptr(out.get("i")),
                                    _lib.ptr(out.get("j")), _lib.ptr(out.get("D")),
                                    _lib.ptr(out.get("d")), _lib.ptr(out.get("S"))))
            else:
                _lib.check(lib.mnl_plan_count_async(*args, None))
                P = None
            if P is not None:
                plan.last_build = (key, P)
            if want_counts:
                out["c"] = torch.empty(self.nat, dtype=torch.int32, **kw)
                _lib.check(lib.mnl_plan_counts(plan.handle, sp, _lib.ptr(out["c"])))
        self.npairs = P
        return out

    @staticmethod
    def _alloc(flags, n, fdt, kw, packed):
        """Output tensors of n rows for the quantities in ``flags``: separate
        tensors, or views of one packed buffer (see ``run``)."""
        torch = _torch()
        parts = [(q, shape, dt) for q, bit, shape, dt in (
            ("D", 4, (n, 3), fdt), ("d", 8, (n,), fdt), ("i", 1, (n,), torch.int32),
            ("j", 2, (n,), torch.int32), ("S", 16, (n, 3), torch.int32)) if flags & bit]
        out = {}
        if packed:
            sizes = [int(np.prod(shape)) * (torch.empty((), dtype=dt).element_size())
                     for _, shape, dt in parts]
            buf = torch.empty(sum(sizes), dtype=torch.uint8, **kw)
            off = 0
            for (q, shape, dt), nb in zip(parts, sizes):
                out[q] = buf[off:off + nb].view(dt).view(shape)
                off += nb
            out["_packed"] = buf
        else:
            for q, shape, dt in parts:
                out[q] = torch.empty(shape, dtype=dt, **kw)
        return out

    @staticmethod
    def _fill_async(lib, plan, sp, capacity, bufs, f32):
        _lib.check(lib.mnl_plan_fill_async(plan.handle, sp, int(capacity),
                                           _lib.ptr(bufs.get("i")), _lib.ptr(bufs.get("j")),
                                           _lib.ptr(bufs.get("D")), _lib.ptr(bufs.get("d")),
                                           _lib.ptr(bufs.get("S")), 1 if f32 else 0))


def _is_f32(out_dtype):
    if out_dtype is None:
        return False
    name = str(getattr(out_dtype, "__name__", out_dtype))
    if "float32" in name:
        return True
    if "float64" in name:
        return False
    raise ValueError(f"out_dtype must be float64 or float32, got {out_dtype!r}")


def _prepare(atoms, positions, cell, pbc, numbers, cell_origin, device):
    """Common front half of every entry point: normalise the arguments and put
    positions on the device. Returns (positions_t, co, ce, inv, pb, nums,
    on_device, want_device_out)."""
    positions, cell, pbc, numbers, cell_origin = _gather(
        atoms, positions, cell, pbc, numbers, cell_origin)
    on_device = _is_on_device(positions)
    want_dev, device_id = _resolve_device(device, on_device)
    if on_device and device is not None and not want_dev:
        raise ValueError("device='cpu' with device-resident positions is not "
                         "supported; move the array to the host first.")
    _lib.require_gpu()
    torch = _torch()
    if on_device:
        pos_t = _device_positions(positions)
        if device_id >= 0 and pos_t.device.index != device_id:
            pos_t = pos_t.to(f"cuda:{device_id}")
        co, ce, inv, pb, nums = _host_metadata(cell, pbc, numbers, cell_origin,
                                               int(pos_t.shape[0]))
    else:
        host_pos = np.ascontiguousarray(positions, dtype=float)
        if host_pos.ndim != 2 or host_pos.shape[1] != 3:
            raise ValueError("positions must have shape (n, 3)")
        co, ce, inv, pb, nums = _host_metadata(cell, pbc, numbers, cell_origin,
                                               len(host_pos), positions=host_pos)
        dev = device_id if device_id >= 0 else torch.cuda.current_device()
        pos_t = torch.from_numpy(host_pos).to(f"cuda:{dev}")
    return pos_t, co, ce, inv, pb, nums, on_device, want_dev


def _export(tensors, want_dev, array_namespace):
    """Hand results over in the requested form (neighbours.py:131-142)."""
    if array_namespace == "dlpack":
        if not want_dev:
            tensors = [t.cpu() for t in tensors]
        return [DLPackTensor(t.__dlpack__(), t.__dlpack_device__()) for t in tensors]
    if array_namespace is None:
        if want_dev:
            return list(tensors)
        return [t.cpu().numpy() for t in tensors]
    if not want_dev:
        tensors = [t.cpu() for t in tensors]
    return [array_namespace.from_dlpack(t) for t in tensors]


def neighbour_list(quantities, atoms=None, cutoff=None, *, positions=None, cell=None,
                   pbc=None, numbers=None, cell_origin=None, device=None,
                   array_namespace=None, out_dtype=None):
    """Compute a neighbour list (neighbours.py:200-288 of the reference).

    quantities: any of ``i`` (first index), ``j`` (second index), ``d``
    (distance), ``D`` (distance vector), ``S`` (cell shift), returned in the
    order given; a single character returns a bare array. ``cutoff`` is a
    float, an element-pair dict or a per-atom radius array. Pairs are sorted by
    ``i`` in the reference's order and ``D == r[j] - r[i] + S @ cell``.
    """
    if cutoff is None:
        raise ValueError("Please provide a value for the cutoff radius.")
    flags = _quantity_flags(quantities)
    f32 = _is_f32(out_dtype)
    pos_t, co, ce, inv, pb, nums, on_device, want_dev = _prepare(
        atoms, positions, cell, pbc, numbers, cell_origin, device)
    rc, nums = _resolve_cutoff(cutoff, nums)
    b = _Build(pos_t, co, ce, inv, pb, rc, nums)
    host_out = not want_dev and array_namespace is None
    out = b.run(flags, f32=f32, packed=host_out)
    if host_out:
        arrays = _packed_to_host(out, quantities)
    else:
        arrays = _export([out[q] for q in quantities], want_dev, array_namespace)
    return arrays[0] if len(quantities) == 1 else tuple(arrays)


# Results up to this size go through pinned host memory (torch's caching
# host allocator); larger ones through one pageable copy, since page-locking
# gigabytes the first time costs more than it saves.
_PINNED_MAX_BYTES = 256 << 20


def _packed_to_host(out, quantities):
    """numpy arrays of a packed result: one transfer of the whole device
    buffer to the host, then a view per quantity."""
    torch = _torch()
    buf = out["_packed"]
    if buf.numel() <= _PINNED_MAX_BYTES:
        host = torch.empty(buf.numel(), dtype=torch.uint8, pin_memory=True)
        host.copy_(buf, non_blocking=True)
        torch.cuda.current_stream(buf.device).synchronize()
    else:
        host = buf.cpu()
    raw = host.numpy()
    base = buf.data_ptr()
    res = []
    for q in quantities:
        t = out[q]
        off = t.data_ptr() - base if t.numel() else 0
        nb = t.numel() * t.element_size()
        dt = np.dtype(str(t.dtype).replace("torch.", ""))
        res.append(raw[off:off + nb].view(dt).reshape(tuple(t.shape)))
    return res


def _batch_geometry(cells, pbc, cutoff, natoms, cell_origins):
    """Per-system host arrays of a batch: atom offsets, origins, cells,
    inv_cell = inv(cell.T) per system (as a single-system call computes it),
    int32 pbc and float cutoffs."""
    natoms = np.asarray(natoms, dtype=np.int64).reshape(-1)
    nsys = natoms.shape[0]
    if nsys < 1:
        raise ValueError("a batch needs at least one system")
    if np.any(natoms < 0):
        raise ValueError("natoms must be non-negative")
    offs = np.zeros(nsys + 1, dtype=np.int64)
    np.cumsum(natoms, out=offs[1:])
    if offs[-1] > np.iinfo(np.int32).max:
        raise ValueError("too many atoms in the batch")
    cells = np.ascontiguousarray(np.broadcast_to(np.asarray(cells, dtype=float), (nsys, 3, 3)))
    invs = np.ascontiguousarray(np.stack([np.linalg.inv(c.T) for c in cells]))
    pb = np.ascontiguousarray(np.broadcast_to(np.asarray(pbc, dtype=bool), (nsys, 3)),
                              dtype=np.int32)
    if isinstance(cutoff, (int, np.integer)) and not isinstance(cutoff, bool):
        raise ValueError("cutoff must be a float (or one float per system)")
    rc = np.ascontiguousarray(np.broadcast_to(np.asarray(cutoff, dtype=float), (nsys,)))
    co = (np.zeros((nsys, 3)) if cell_origins is None else
          np.ascontiguousarray(np.broadcast_to(np.asarray(cell_origins, dtype=float),
                                               (nsys, 3))))
    return nsys, offs.astype(np.int32), co, cells, invs, pb, rc


def neighbour_list_batch(quantities, positions, cells, pbc, cutoff, natoms, *,
                         cell_origins=None, device=None, array_namespace=None,
                         out_dtype=None):
    """Neighbour lists of many independent systems in one build.

    No counterpart in the reference, whose API builds one system per call:
    every system's pairs are exactly those ``neighbour_list`` gives for it
    alone (i and j local to the system, same order, same bits), concatenated
    in system order.

    positions: (sum(natoms), 3) float64, systems concatenated (host or
    device); cells: (nsys, 3, 3) lattice vectors as rows; pbc: (nsys, 3),
    (3,) or a bool; cutoff: a float or one float per system; natoms: atoms
    per system. Returns the requested quantities followed by
    ``system_pair_offsets`` (nsys + 1,) int64: system s owns pairs
    ``[offsets[s], offsets[s + 1])``.
    """
    flags = _quantity_flags(quantities)
    f32 = _is_f32(out_dtype)
    nsys, offs, co, ce, inv, pb, rc = _batch_geometry(cells, pbc, cutoff, natoms,
                                                      cell_origins)
    on_device = _is_on_device(positions)
    want_dev, device_id = _resolve_device(device, on_device)
    _lib.require_gpu()
    torch = _torch()
    if on_device:
        pos_t = _device_positions(positions)
        if device_id >= 0 and pos_t.device.index != device_id:
            pos_t = pos_t.to(f"cuda:{device_id}")
    else:
        host_pos = np.ascontiguousarray(positions, dtype=float)
        if host_pos.ndim != 2 or host_pos.shape[1] != 3:
            raise ValueError("positions must have shape (n, 3)")
        dev = device_id if device_id >= 0 else torch.cuda.current_device()
        pos_t = torch.from_numpy(host_pos).to(f"cuda:{dev}")
    if pos_t.shape[0] != int(offs[-1]):
        raise ValueError("positions must hold sum(natoms) rows")
    lib = _lib.load()
    idx = pos_t.device.index if pos_t.device.index is not None else torch.cuda.current_device()
    plan = _lib.plan_for(idx)
    out = {}
    kw = dict(device=pos_t.device)
    with plan.lock, torch.cuda.device(idx):
        sp = _lib.ctypes.c_void_p(torch.cuda.current_stream(idx).cuda_stream)
        _lib.check(lib.mnl_plan_set_batch(plan.handle, sp, nsys, _lib.ptr(offs),
                                          _lib.ptr(co), _lib.ptr(ce), _lib.ptr(inv),
                                          _lib.ptr(pb), _lib.ptr(rc)))
        npairs = _lib.ctypes.c_int64(0)
        _lib.check(lib.mnl_plan_batch_count(plan.handle, sp, _lib.ptr(pos_t),
                                            _lib.ctypes.byref(npairs)))
        P = int(npairs.value)
        fdt = torch.float32 if f32 else torch.float64
        shapes = {"i": ((P,), torch.int32), "j": ((P,), torch.int32), "D": ((P, 3), fdt),
                  "d": ((P,), fdt), "S": ((P, 3), torch.int32)}
        for q in "ijDdS":
            if flags & _lib.QUANTITY_BITS[q]:
                out[q] = torch.empty(shapes[q][0], dtype=shapes[q][1], **kw)
        if flags:
            fill = lib.mnl_plan_fill_f32 if f32 else lib.mnl_plan_fill
            _lib.check(fill(plan.handle, sp, _lib.ptr(out.get("i")), _lib.ptr(out.get("j")),
                            _lib.ptr(out.get("D")), _lib.ptr(out.get("d")),
                            _lib.ptr(out.get("S"))))
        sys_off = torch.empty(nsys + 1, dtype=torch.int64, **kw)
        _lib.check(lib.mnl_plan_system_offsets(plan.handle, sp, _lib.ptr(sys_off)))
    tensors = [out[q] for q in quantities] + [sys_off]
    return tuple(_export(tensors, want_dev, array_namespace))


def neighbour_matrix(atoms=None, cutoff=None, max_neighbours=None, *, positions=None,
                     cell=None, pbc=None, numbers=None, cell_origin=None, device=None,
                     array_namespace=None):
    """Dense fixed-capacity neighbour list ``(idx, dist, count)``
    (neighbours.py:291-357 of the reference).

    ``idx`` (n, max_neighbours) int32 holds each atom's neighbours in the
    pair list's order (unused slots 0), ``dist`` (n, max_neighbours, 3) the
    distance vectors ``D`` (unused slots 0), ``count`` (n,) the true
    neighbour counts. Raises ``ValueError`` if any atom has more than
    ``max_neighbours`` neighbours. Built in one traversal on the GPU (no
    pair list)."""
    if cutoff is None:
        raise ValueError("Please provide a value for the cutoff radius.")
    if max_neighbours is None:
        raise ValueError("Please provide max_neighbours (the row capacity).")
    K = int(max_neighbours)
    pos_t, co, ce, inv, pb, nums, on_device, want_dev = _prepare(
        atoms, positions, cell, pbc, numbers, cell_origin, device)
    rc, nums = _resolve_cutoff(cutoff, nums)
    b = _Build(pos_t, co, ce, inv, pb, rc, nums)
    torch = _torch()
    lib = _lib.load()
    n = b.nat
    idx = b.dev.index if b.dev.index is not None else torch.cuda.current_device()
    plan = _lib.plan_for(idx)
    kw = dict(device=b.dev)
    m_idx = torch.empty((n, K), dtype=torch.int32, **kw)
    m_dist = torch.empty((n, K, 3), dtype=torch.float64, **kw)
    m_cnt = torch.empty(n, dtype=torch.int32, **kw)
    overflow = _lib.ctypes.c_int32(0)
    with plan.lock, torch.cuda.device(idx):
        sp = _lib.ctypes.c_void_p(torch.cuda.current_stream(idx).cuda_stream)
        _lib.check(lib.mnl_plan_matrix(
            plan.handle, sp, _lib.ptr(b.co), _lib.ptr(b.ce), _lib.ptr(b.inv), _lib.ptr(b.pbc),
            n, _lib.ptr(b.pos), b.rc, _lib.ptr(b.per_atom_t), _lib.ptr(b.per_type_sq), b.nc,
            _lib.ptr(b.types_t), -1, K, _lib.ptr(m_idx), _lib.ptr(m_dist), _lib.ptr(m_cnt),
            _lib.ctypes.byref(overflow)))
    if overflow.value:
        raise ValueError(
            f"max_neighbours={K} is too small: some atom has more neighbours than that. "
            "Retry with a larger max_neighbours.")
    return tuple(_export([m_idx, m_dist, m_cnt], want_dev, array_namespace))


def coordination(atoms=None, cutoff=None, *, positions=None, cell=None, pbc=None,
                 numbers=None, cell_origin=None, device=None, array_namespace=None):
    """Neighbours per atom within ``cutoff`` (neighbours.py:373-422), from the
    count pass only: the pair list is never materialised."""
    if cutoff is None:
        raise ValueError("Please provide a value for the cutoff radius.")
    pos_t, co, ce, inv, pb, nums, on_device, want_dev = _prepare(
        atoms, positions, cell, pbc, numbers, cell_origin, device)
    rc, nums = _resolve_cutoff(cutoff, nums)
    b = _Build(pos_t, co, ce, inv, pb, rc, nums)
    out = b.run(0, want_counts=True, need_total=False)
    counts = out["c"]
    if not want_dev and array_namespace is None:
        return counts.cpu().numpy().astype(np.int64)
    return _export([counts], want_dev, array_namespace)[0]


def lennard_jones(atoms=None, cutoff=None, *, epsilon=1.0, sigma=1.0, positions=None,
                  cell=None, pbc=None, cell_origin=None, device=None, array_namespace=None):
    """Lennard-Jones ``(energy, forces, virial)`` over the neighbour list's
    pairs (r < cutoff, periodic images included), computed on the GPU
    straight from the cell grid: the pair list is never written.

    phi(r) = 4 epsilon ((sigma/r)^12 - (sigma/r)^6), unshifted; forces (n, 3)
    f_i = sum_j coef D_ij with coef = -24 epsilon (2 (sigma/r)^12 -
    (sigma/r)^6) / r^2 (the force loop of the reference's
    examples/lj_langevin/lj_langevin_cpu.cc:76-90, there with epsilon =
    sigma = 1); virial (3, 3) = sum over pairs of r_ij (x) f_ij. Host input
    gives a float and numpy arrays, device input device tensors."""
    if cutoff is None:
        raise ValueError("Please provide a value for the cutoff radius.")
    if not isinstance(cutoff, (float, int)) or isinstance(cutoff, bool):
        raise TypeError("lennard_jones takes a single float cutoff.")
    pos_t, co, ce, inv, pb, nums, on_device, want_dev = _prepare(
        atoms, positions, cell, pbc, None, cell_origin, device)
    b = _Build(pos_t, co, ce, inv, pb, float(cutoff), nums)
    torch = _torch()
    lib = _lib.load()
    n = b.nat
    idx = b.dev.index if b.dev.index is not None else torch.cuda.current_device()
    plan = _lib.plan_for(idx)
    kw = dict(dtype=torch.float64, device=b.dev)
    forces = torch.empty((n, 3), **kw)
    energy = torch.empty(1, **kw)
    virial = torch.empty((3, 3), **kw)
    with plan.lock, torch.cuda.device(idx):
        sp = _lib.ctypes.c_void_p(torch.cuda.current_stream(idx).cuda_stream)
        _lib.check(lib.mnl_plan_lj(
            plan.handle, sp, _lib.ptr(b.co), _lib.ptr(b.ce), _lib.ptr(b.inv), _lib.ptr(b.pbc),
            n, _lib.ptr(b.pos), b.rc, float(epsilon), float(sigma), _lib.ptr(forces),
            _lib.ptr(energy), _lib.ptr(virial)))
    if not want_dev and array_namespace is None:
        return float(energy.item()), forces.cpu().numpy(), virial.cpu().numpy()
    e, f, w = _export([energy.reshape(()), forces, virial], want_dev, array_namespace)
    return e, f, w


def _tune_plan(dev):
    torch = _torch()
    lib = _lib.load()
    idx = dev.index if dev.index is not None else torch.cuda.current_device()
    plan = _lib.plan_for(idx)
    split, chunks, ms = _lib.ctypes.c_int32(0), _lib.ctypes.c_int32(0), _lib.ctypes.c_float(0)
    with plan.lock, torch.cuda.device(idx):
        sp = _lib.ctypes.c_void_p(torch.cuda.current_stream(idx).cuda_stream)
        _lib.check(lib.mnl_plan_autotune(plan.handle, sp, _lib.ctypes.byref(split),
                                         _lib.ctypes.byref(chunks), _lib.ctypes.byref(ms)))
    return {"split": split.value, "chunks": chunks.value, "ms": ms.value}


def autotune(atoms=None, cutoff=None, *, positions=None, cell=None, pbc=None, numbers=None,
             cell_origin=None, device=None):
    """Tune the pair kernels' launch shape for this system's configuration
    class (atom count, cells, grid backend, cutoff form) on the GPU: the
    fastest candidate shape is kept for the process and used by every later
    build of the same class. Outputs never depend on the shape. Returns
    ``{"split", "chunks", "ms"}``. No counterpart in the reference."""
    if cutoff is None:
        raise ValueError("Please provide a value for the cutoff radius.")
    pos_t, co, ce, inv, pb, nums, on_device, want_dev = _prepare(
        atoms, positions, cell, pbc, numbers, cell_origin, device)
    rc, nums = _resolve_cutoff(cutoff, nums)
    b = _Build(pos_t, co, ce, inv, pb, rc, nums)
    b.run(0)
    return _tune_plan(b.dev)


def autotune_batch(positions, cells, pbc, cutoff, natoms, *, cell_origins=None, device=None):
    """``autotune`` for the configuration class of a batch (arguments as
    ``neighbour_list_batch``)."""
    nsys, offs, co, ce, inv, pb, rc = _batch_geometry(cells, pbc, cutoff, natoms, cell_origins)
    on_device = _is_on_device(positions)
    want_dev, device_id = _resolve_device(device, on_device)
    _lib.require_gpu()
    torch = _torch()
    if on_device:
        pos_t = _device_positions(positions)
    else:
        dev = device_id if device_id >= 0 else torch.cuda.current_device()
        pos_t = torch.from_numpy(np.ascontiguousarray(positions, dtype=float)).to(f"cuda:{dev}")
    if pos_t.shape[0] != int(offs[-1]):
        raise ValueError("positions must hold sum(natoms) rows")
    lib = _lib.load()
    idx = pos_t.device.index if pos_t.device.index is not None else torch.cuda.current_device()
    plan = _lib.plan_for(idx)
    with plan.lock, torch.cuda.device(idx):
        sp = _lib.ctypes.c_void_p(torch.cuda.current_stream(idx).cuda_stream)
        _lib.check(lib.mnl_plan_set_batch(plan.handle, sp, nsys, _lib.ptr(offs), _lib.ptr(co),
                                          _lib.ptr(ce), _lib.ptr(inv), _lib.ptr(pb),
                                          _lib.ptr(rc)))
        npairs = _lib.ctypes.c_int64(0)
        _lib.check(lib.mnl_plan_batch_count(plan.handle, sp, _lib.ptr(pos_t),
                                            _lib.ctypes.byref(npairs)))
    return _tune_plan(pos_t.device)


def _int32_on_device(x):
    """(tensor on the GPU, was_device) for an int32 index array."""
    torch = _torch()
    if _is_on_device(x):
        t = x if isinstance(x, torch.Tensor) else torch.from_dlpack(x)
        return t.to(torch.int32).contiguous(), True
    a = np.ascontiguousarray(x, dtype=np.int32).reshape(-1)
    return torch.from_numpy(a).to(f"cuda:{torch.cuda.current_device()}"), False


def first_neighbours(n, i):
    """Row-start index of each atom in an i-sorted pair list
    (first_neighbours.cc:15-40): seed (n + 1,) int32; -1 for atoms before the
    first populated row, nn after the last. Computed on the GPU; host input
    gives a host result."""
    _lib.require_gpu()
    torch = _torch()
    lib = _lib.load()
    n = int(n)
    it, dev_in = _int32_on_device(i)
    seed = torch.empty(n + 1, dtype=torch.int32, device=it.device)
    with torch.cuda.device(it.device):
        sp = _lib.ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(lib.mnl_first_neighbours(sp, n, int(it.numel()), _lib.ptr(it),
                                            _lib.ptr(seed)))
    return seed if dev_in else seed.cpu().numpy()


def get_jump_indicies(sorted_array):
    """Positions where a sorted array changes value, plus both ends
    (first_neighbours.cc:42-55). Computed on the GPU."""
    _lib.require_gpu()
    torch = _torch()
    lib = _lib.load()
    it, dev_in = _int32_on_device(sorted_array)
    plan = _lib.plan_for(it.device.index)
    nd = _lib.ctypes.c_int32(0)
    with plan.lock, torch.cuda.device(it.device):
        sp = _lib.ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(lib.mnl_plan_count_distinct(plan.handle, sp, int(it.numel()), _lib.ptr(it),
                                               _lib.ctypes.byref(nd)))
    out = first_neighbours(int(nd.value), it)
    return out if dev_in else out.cpu().numpy()


def triplet_list(first_neighbours, abs_dr_p=None, cutoff=None):
    """Triplets (ij, ik) of pairs sharing the first atom (triplet_list.cc:15-34,
    neighbours.py:360-370 of the reference): computed on the GPU; host
    input gives host arrays."""
    if (abs_dr_p is None) != (cutoff is None):
        raise TypeError("Cutoff and distances must be specified together.")
    if abs_dr_p is not None and not isinstance(cutoff, float):
        raise NotImplementedError("Cutoff must be a single float.")
    _lib.require_gpu()
    torch = _torch()
    lib = _lib.load()
    ft, dev_in = _int32_on_device(first_neighbours)
    d = None
    if abs_dr_p is not None:
        if _is_on_device(abs_dr_p):
            d = abs_dr_p if isinstance(abs_dr_p, torch.Tensor) else torch.from_dlpack(abs_dr_p)
            d = d.to(device=ft.device, dtype=torch.float64).contiguous()
        else:
            d = torch.from_numpy(np.ascontiguousarray(abs_dr_p, dtype=np.float64)).to(ft.device)
    nf = int(ft.numel())
    nn = int(d.numel()) if d is not None else (int(ft[-1].item()) if nf else 0)
    nn = max(nn, 0)
    plan = _lib.plan_for(ft.device.index)
    nt = _lib.ctypes.c_int64(0)
    with plan.lock, torch.cuda.device(ft.device):
        sp = _lib.ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(lib.mnl_plan_triplets_count(plan.handle, sp, nf, _lib.ptr(ft), nn,
                                               _lib.ptr(d), float(cutoff or 0.0),
                                               _lib.ctypes.byref(nt)))
        ij = torch.empty(int(nt.value), dtype=torch.int32, device=ft.device)
        ik = torch.empty(int(nt.value), dtype=torch.int32, device=ft.device)
        _lib.check(lib.mnl_plan_triplets_fill(plan.handle, sp, _lib.ptr(ij), _lib.ptr(ik)))
    if dev_in:
        return ij, ik
    return ij.cpu().numpy(), ik.cpu().numpy()


_lock = threading.Lock()
