"""The deterministic backward (mnl_plan_pair_backward with deterministic=1).

Its sums run in a fixed order, which numpy restates exactly:
  g_p = gD_p + (gd_p / d_p) D_p (the d term dropped where d_p = 0),
  dL/dr[a] = W(pairs with j = a, in ascending pair index) - W(the row of a),
  where W(list) gives lane l (0..63) the in-order sum, from 0.0, of entries
  l, l + 64, ... and combines the 64 lane sums by the halving tree
  v[l] += v[l + h], h = 32, 16, ..., 1 (the result is v[0]),
  dL/dcell: per block b of min(1024, ceil(P/256)) blocks, thread t adds
  pairs b*256 + t, + stride, ... in order, the 256 threads are combined in a
  halving tree; then thread t of one block adds the block results t, t + 256,
  ... in order and the same tree combines the 256 threads (periodic
  directions of S only).
So the GPU result must equal the restatement bit for bit, on every call;
it must also agree with the atomic (default) backward to rounding."""

import ctypes

import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu


def _g(D, d, gD, gd):
    g = np.zeros_like(D) if gD is None else gD.astype(np.float64).copy()
    if gd is not None:
        dd = d.astype(np.float64)
        f = np.zeros_like(dd)
        np.divide(gd.astype(np.float64), dd, out=f, where=dd > 0.0)
        g = np.where((dd > 0.0)[:, None], g + f[:, None] * D.astype(np.float64), g)
    return g


def _wave_sums(owner, vals, nat):
    """W(list) of every atom at once. owner[k] is the atom of entry k; the
    entries are grouped by atom, each group in summation order. Lane l of
    atom a adds its entries l, l + 64, ... in order from 0.0; the 64 lane
    sums are combined by the fixed halving tree."""
    start = np.searchsorted(owner, np.arange(nat + 1))
    rank = np.arange(len(owner)) - start[owner] if len(owner) else np.zeros(0, np.int64)
    out = np.zeros((nat, 3))
    chunk = 65536
    for a0 in range(0, nat, chunk):
        a1 = min(nat, a0 + chunk)
        sl = slice(start[a0], start[a1])
        o, r, v = owner[sl] - a0, rank[sl], vals[sl]
        lanes = np.zeros((a1 - a0, 64, 3))
        for k in range(int(r.max()) // 64 + 1 if len(r) else 0):
            m = (r // 64) == k            # at most one entry per (atom, lane)
            lanes[o[m], r[m] % 64] = lanes[o[m], r[m] % 64] + v[m]
        h = 32
        while h > 0:
            lanes[:, :h] = lanes[:, :h] + lanes[:, h:2 * h]
            h //= 2
        out[a0:a1] = lanes[:, 0]
    return out


def _restated(i, j, S, D, d, gD, gd, nat, pbc):
    g = _g(D, d, gD, gd)
    by_j = np.argsort(j, kind="stable")          # ascending pair index within each j
    gpos = _wave_sums(j[by_j], g[by_j], nat) - _wave_sums(i, g, nat)
    P = len(i)
    nb = min(1024, (P + 255) // 256)
    stride = nb * 256
    s = np.where(np.asarray(pbc, bool)[None, :], S.astype(np.float64), 0.0)
    terms = (s[:, :, None] * g[:, None, :]).reshape(P, 9)
    pad = (-P) % stride
    t = np.concatenate([terms, np.zeros((pad, 9))]).reshape(-1, nb, 256, 9)
    acc = np.zeros((nb, 256, 9))
    for k in range(t.shape[0]):          # each thread's pairs in order
        acc = acc + t[k]
    h = 128
    while h > 0:                         # the halving tree over threads
        acc[:, :h] = acc[:, :h] + acc[:, h:2 * h]
        h //= 2
    part = np.concatenate([acc[:, 0], np.zeros(((-nb) % 256, 9))]).reshape(-1, 256, 9)
    fin = np.zeros((256, 9))
    for k in range(part.shape[0]):       # thread t: block results t, t + 256, ...
        fin = fin + part[k]
    h = 128
    while h > 0:
        fin[:h] = fin[:h] + fin[h:2 * h]
        h //= 2
    return gpos, fin[0].reshape(3, 3)


def _backward(lib, _lib, torch, i, j, S, D, d, gD, gd, nat, pbc, det):
    plan = _lib.plan_for(0)
    gpos = torch.full((nat, 3), np.nan, dtype=torch.float64, device="cuda:0")
    gcell = torch.full((3, 3), np.nan, dtype=torch.float64, device="cuda:0")
    pb = np.asarray(pbc, np.int32)
    f32 = D.dtype == torch.float32
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(lib.mnl_plan_pair_backward(plan.handle, sp, len(i), nat, _lib.ptr(i), _lib.ptr(j),
                                          _lib.ptr(S), _lib.ptr(D), _lib.ptr(d), _lib.ptr(gD),
                                          _lib.ptr(gd), 1 if f32 else 0, _lib.ptr(pb),
                                          1 if det else 0, _lib.ptr(gpos), _lib.ptr(gcell)))
    torch.cuda.synchronize()
    return gpos.cpu().numpy(), gcell.cpu().numpy()


@pytest.mark.parametrize("case", ["random_TTF", "triclinic_TTT_f32", "tiny_cell", "D_only"])
def test_deterministic_backward_is_the_fixed_order_sum(gpu, case):
    import torch

    from matscipy_neighbours_amd import _lib, neighbour_list
    lib = _lib.load()
    rng = np.random.default_rng({"random_TTF": 1, "triclinic_TTT_f32": 2, "tiny_cell": 3,
                                 "D_only": 4}[case])
    if case == "tiny_cell":
        cell = np.array([[2.0, 0.0, 0.0], [0.3, 1.9, 0.0], [0.2, -0.1, 2.2]])
        n, rc, pbc = 7, 4.1, (True, True, True)
    else:
        cell = np.diag([18.0, 16.0, 17.0]) + rng.uniform(-1.0, 1.0, (3, 3))
        n, rc = 3000, 3.3
        pbc = (True, True, False) if case == "random_TTF" else (True, True, True)
    pos = rng.uniform(-0.2, 1.2, (n, 3)) @ cell
    out_dtype = torch.float32 if case == "triclinic_TTT_f32" else None
    t = torch.tensor(pos, device="cuda:0")
    i, j, D, d, S = neighbour_list("ijDdS", positions=t, cell=cell, pbc=pbc, cutoff=rc,
                                   out_dtype=out_dtype)
    P = len(i)
    print(f"{case}: {n} atoms, {P} pairs", flush=True)
    gD = torch.tensor(rng.normal(size=(P, 3)), dtype=D.dtype, device="cuda:0")
    gd = None if case == "D_only" else torch.tensor(rng.normal(size=P), dtype=D.dtype,
                                                    device="cuda:0")
    a = _backward(lib, _lib, torch, i, j, S, D, d, gD, gd, n, pbc, True)
    b = _backward(lib, _lib, torch, i, j, S, D, d, gD, gd, n, pbc, True)
    assert np.array_equal(bits(a[0]), bits(b[0])) and np.array_equal(bits(a[1]), bits(b[1]))
    want = _restated(i.cpu().numpy(), j.cpu().numpy(), S.cpu().numpy(), D.cpu().numpy(),
                     d.cpu().numpy(), gD.cpu().numpy(), None if gd is None else gd.cpu().numpy(),
                     n, pbc)
    assert np.array_equal(bits(a[0]), bits(want[0]))
    assert np.array_equal(bits(a[1]), bits(want[1]))
    atomic = _backward(lib, _lib, torch, i, j, S, D, d, gD, gd, n, pbc, False)
    scale = max(float(np.abs(want[0]).max()), 1e-300)
    assert float(np.abs(atomic[0] - a[0]).max()) <= 1e-12 * scale
    assert np.allclose(atomic[1], a[1], rtol=1e-12, atol=1e-12 * float(np.abs(a[1]).max()))


def test_deterministic_autograd_on_bcc(gpu):
    """The 131,072-atom sheared bcc at rc 6: two deterministic backward calls
    give the same bits, and agree with the default backward."""
    import torch

    from matscipy_neighbours_amd import workloads
    from matscipy_neighbours_amd.autograd import neighbour_list_autograd
    w = workloads.bcc_sheared()
    grads = []
    for det in (True, True, False):
        pos = torch.tensor(w["positions"], device="cuda:0", requires_grad=True)
        cell = torch.tensor(w["cell"], device="cuda:0", requires_grad=True)
        d = neighbour_list_autograd("d", pos, cell, True, 6.0, deterministic=det)
        ((1.0 - d / 6.0) ** 3).sum().backward()
        grads.append((pos.grad.cpu().numpy(), cell.grad.cpu().numpy()))
    assert np.array_equal(bits(grads[0][0]), bits(grads[1][0]))
    assert np.array_equal(bits(grads[0][1]), bits(grads[1][1]))
    scale = float(np.abs(grads[2][0]).max())
    assert float(np.abs(grads[0][0] - grads[2][0]).max()) <= 1e-12 * scale
    assert np.allclose(grads[0][1], grads[2][1], rtol=1e-11)


def test_deterministic_backward_three_radix_passes(gpu):
    """1,200,000 atoms (keys above 2**20, so three 7-bit radix passes) at
    about two neighbours each: bit-exact to the restatement."""
    import torch

    from matscipy_neighbours_amd import _lib, neighbour_list
    lib = _lib.load()
    rng = np.random.default_rng(9)
    n, rc, pbc = 1_200_000, 3.0, (True, False, True)
    cell = np.diag([400.0, 400.0, 400.0])
    pos = rng.uniform(0.0, 1.0, (n, 3)) @ cell
    t = torch.tensor(pos, device="cuda:0")
    i, j, D, d, S = neighbour_list("ijDdS", positions=t, cell=cell, pbc=pbc, cutoff=rc)
    P = len(i)
    print(f"{n} atoms, {P} pairs", flush=True)
    gD = torch.tensor(rng.normal(size=(P, 3)), device="cuda:0")
    gd = torch.tensor(rng.normal(size=P), device="cuda:0")
    a = _backward(lib, _lib, torch, i, j, S, D, d, gD, gd, n, pbc, True)
    want = _restated(i.cpu().numpy(), j.cpu().numpy(), S.cpu().numpy(), D.cpu().numpy(),
                     d.cpu().numpy(), gD.cpu().numpy(), gd.cpu().numpy(), n, pbc)
    assert np.array_equal(bits(a[0]), bits(want[0]))
    assert np.array_equal(bits(a[1]), bits(want[1]))
