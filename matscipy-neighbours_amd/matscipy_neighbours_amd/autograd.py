"""Differentiable neighbour lists (PyTorch autograd).

No counterpart in the reference, which has no derivatives. The forward is
the engine's ordinary build (the pair set decided in float64, D and d
written in float64 or float32); the backward is the engine's HIP kernel for
D_p = r[j] - r[i] + S_p @ cell and d_p = |D_p|:

    g_p           = dL/dD_p + dL/dd_p * D_p / d_p
    dL/dr[j]     += g_p,   dL/dr[i] -= g_p
    dL/dcell[a,b] = sum_p S_pa g_pb   (periodic directions a only: along a
                                       non-periodic one D = r_j - r_i)

The pair set (i, j, S) is piecewise constant in the positions and the cell,
so it carries no gradient, exactly as for a list built on the CPU.
"""

import ctypes

import numpy as np

from . import _lib
from .neighbours import _Build, _is_f32, _quantity_flags

__all__ = ["neighbour_list_autograd"]


def _function():
    import torch

    class NeighbourListFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, positions, cell, pbc, cutoff, f32, cell_origin, deterministic):
            pos = positions.detach()
            if pos.dtype != torch.float64 or not pos.is_contiguous():
                pos = pos.to(torch.float64).contiguous()
            ce = cell.detach().to("cpu", torch.float64).numpy().copy()
            co = (np.zeros(3) if cell_origin is None else
                  np.ascontiguousarray(cell_origin, dtype=np.float64))
            inv = np.ascontiguousarray(np.linalg.inv(ce.T))
            b = _Build(pos, co, ce, inv, pbc, cutoff, None)
            out = b.run(31, f32=f32)
            i, j, D, d, S = out["i"], out["j"], out["D"], out["d"], out["S"]
            ctx.save_for_backward(i, j, S, D, d)
            ctx.pbc = np.ascontiguousarray(pbc, dtype=np.int32)
            ctx.deterministic = bool(deterministic)
            ctx.nat = int(pos.shape[0])
            ctx.f32 = f32
            ctx.pos_dtype = positions.dtype
            ctx.cell_dtype = cell.dtype
            ctx.mark_non_differentiable(i, j, S)
            ctx.set_materialize_grads(False)  # unused outputs arrive as None
            return i, j, D, d, S

        @staticmethod
        def backward(ctx, gi, gj, gD, gd, gS):
            i, j, S, D, d = ctx.saved_tensors
            dev = D.device
            need_pos, need_cell = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
            if gD is None and gd is None:
                return None, None, None, None, None, None, None
            fdt = torch.float32 if ctx.f32 else torch.float64
            gD = None if gD is None else gD.to(fdt).contiguous()
            gd = None if gd is None else gd.to(fdt).contiguous()
            grad_pos = torch.empty((ctx.nat, 3), dtype=torch.float64, device=dev)
            grad_cell = torch.empty((3, 3), dtype=torch.float64, device=dev)
            lib = _lib.load()
            idx = dev.index if dev.index is not None else torch.cuda.current_device()
            plan = _lib.plan_for(idx)
            with plan.lock, torch.cuda.device(idx):
                sp = ctypes.c_void_p(torch.cuda.current_stream(idx).cuda_stream)
                _lib.check(lib.mnl_plan_pair_backward(
                    plan.handle, sp, int(i.shape[0]), ctx.nat, _lib.ptr(i), _lib.ptr(j),
                    _lib.ptr(S), _lib.ptr(D), _lib.ptr(d), _lib.ptr(gD), _lib.ptr(gd),
                    1 if ctx.f32 else 0, _lib.ptr(ctx.pbc), 1 if ctx.deterministic else 0,
                    _lib.ptr(grad_pos), _lib.ptr(grad_cell)))
            gp = grad_pos.to(ctx.pos_dtype) if need_pos else None
            gc = grad_cell.to(ctx.cell_dtype) if need_cell else None
            return gp, gc, None, None, None, None, None

    return NeighbourListFunction


_FN = None


def neighbour_list_autograd(quantities, positions, cell, pbc, cutoff, *, cell_origin=None,
                            out_dtype=None, deterministic=False):
    """Neighbour list whose D and d are differentiable with respect to
    ``positions`` ((n, 3) float64 torch tensor on the GPU) and ``cell``
    ((3, 3) torch tensor, lattice vectors as rows). ``cutoff`` is a float;
    ``quantities`` as for ``neighbour_list`` (results in that order, torch
    tensors on the positions' device). ``out_dtype=torch.float32`` gives
    float32 D and d (pairs still decided in float64). ``deterministic=True``
    makes the position gradient the same bits on every call (fixed-order
    sums instead of float64 atomics; slower, see mnl_plan_pair_backward)."""
    global _FN
    import torch
    if not isinstance(positions, torch.Tensor) or positions.device.type != "cuda":
        raise TypeError("positions must be a torch tensor on a GPU device")
    if not isinstance(cutoff, float):
        raise TypeError("neighbour_list_autograd takes a float cutoff")
    if not isinstance(cell, torch.Tensor):
        cell = torch.as_tensor(np.asarray(cell, dtype=np.float64))
    _quantity_flags(quantities)
    _lib.require_gpu()
    if _FN is None:
        _FN = _function()
    pbc = np.ascontiguousarray(np.broadcast_to(np.asarray(pbc, dtype=bool), (3,)))
    i, j, D, d, S = _FN.apply(positions, cell, pbc, cutoff, _is_f32(out_dtype), cell_origin,
                              bool(deterministic))
    res = dict(i=i, j=j, D=D, d=d, S=S)
    out = tuple(res[q] for q in quantities)
    return out[0] if len(out) == 1 else out
