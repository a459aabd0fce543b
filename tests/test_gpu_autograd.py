"""The differentiable list: the HIP backward of D = r[j] - r[i] + S @ cell and
d = |D| against

* torch autograd of that expression, evaluated on the engine's own (i, j, S)
  (the reference defines D this way; tests/test_extension.py:50-92 of the
  reference checks exactly this identity on its outputs);
* float64 central finite differences (h = 1e-6 A) of the upstream loss
  sum (1 - d/rc)^3, rebuilding the list at every displaced geometry;
* torch.autograd.gradcheck on a small triclinic system;

on the 131,072-atom sheared bcc workload (rc 6 A) in float64 and float32.
Tolerances: the backward and the autograd reference sum the same terms in a
different order (the j-side uses float64 atomics), so they agree to ~1e-12
relative; finite differences carry a rounding error of ~eps*|L|/h."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RC = 6.0


@pytest.fixture(scope="module")
def bcc():
    from matscipy_neighbours_amd import workloads
    return workloads.bcc_sheared()


def _loss(d):
    return ((1.0 - d / RC) ** 3).sum()


def _setup(w, dtype=None):
    import torch
    pos = torch.tensor(w["positions"], dtype=torch.float64, device="cuda:0",
                       requires_grad=True)
    cell = torch.tensor(w["cell"], dtype=torch.float64, device="cuda:0", requires_grad=True)
    return pos, cell


def _autograd_reference(pos, cell, i, j, S, weights):
    """Gradients of sum (1 - d/rc)^3 + sum w . D with D = r[j] - r[i] + S @ cell,
    by torch autograd in float64."""
    import torch
    p = pos.detach().clone().requires_grad_(True)
    c = cell.detach().clone().requires_grad_(True)
    D = p[j.long()] - p[i.long()] + S.to(torch.float64) @ c
    d = torch.linalg.norm(D, dim=1)
    L = _loss(d) + (weights * D).sum()
    L.backward()
    return D.detach(), d.detach(), p.grad, c.grad


def _rel_err(a, b):
    return float((a - b).abs().max() / max(float(b.abs().max()), 1e-300))


def test_backward_matches_torch_autograd(gpu, bcc):
    import torch

    from matscipy_neighbours_amd.autograd import neighbour_list_autograd
    pos, cell = _setup(bcc)
    i, j, D, d, S = neighbour_list_autograd("ijDdS", pos, cell, True, RC)
    P = len(i)
    print(f"bcc: {pos.shape[0]} atoms, {P} pairs", flush=True)
    assert 8.0e6 < P < 9.0e6  # bcc shells within 6 A: 8+6+12+24+8+6 = 64 per atom, + jitter
    w = torch.randn((P, 3), dtype=torch.float64, device="cuda:0", generator=torch.Generator(
        device="cuda:0").manual_seed(3))
    L = _loss(d) + (w * D).sum()
    L.backward()
    D_ref, d_ref, gp_ref, gc_ref = _autograd_reference(pos, cell, i, j, S, w)
    # forward: the engine's D equals r[j] - r[i] + S @ cell to rounding
    assert _rel_err(D.detach(), D_ref) < 1e-12
    assert _rel_err(d.detach(), d_ref) < 1e-12
    assert _rel_err(pos.grad, gp_ref) < 1e-10
    assert _rel_err(cell.grad, gc_ref) < 1e-10


def test_backward_shuffled_atoms(gpu, bcc):
    """The bcc atoms in random order: the rows of a block's tile are atoms
    far apart, so its LDS table of j-side sums fills (every slot used, then
    pairs adding straight to global memory), and the sorted flush adds a full
    table. Gradients against torch autograd as above."""
    import torch

    from matscipy_neighbours_amd.autograd import neighbour_list_autograd
    perm = np.random.default_rng(11).permutation(len(bcc["positions"]))
    shuffled = {"positions": bcc["positions"][perm], "cell": bcc["cell"]}
    pos, cell = _setup(shuffled)
    i, j, D, d, S = neighbour_list_autograd("ijDdS", pos, cell, True, RC)
    P = len(i)
    print(f"shuffled bcc: {pos.shape[0]} atoms, {P} pairs", flush=True)
    assert 8.0e6 < P < 9.0e6
    w = torch.randn((P, 3), dtype=torch.float64, device="cuda:0", generator=torch.Generator(
        device="cuda:0").manual_seed(4))
    L = _loss(d) + (w * D).sum()
    L.backward()
    _, _, gp_ref, gc_ref = _autograd_reference(pos, cell, i, j, S, w)
    assert _rel_err(pos.grad, gp_ref) < 1e-10
    assert _rel_err(cell.grad, gc_ref) < 1e-10


def test_backward_d_only_and_D_only(gpu, bcc):
    """Each incoming gradient alone (the other is None)."""
    import torch

    from matscipy_neighbours_amd.autograd import neighbour_list_autograd
    pos, cell = _setup(bcc)
    i, j, d, S = neighbour_list_autograd("ijdS", pos, cell, True, RC)
    _loss(d).backward()
    zero = torch.zeros((len(i), 3), dtype=torch.float64, device="cuda:0")
    _, _, gp_ref, gc_ref = _autograd_reference(pos, cell, i, j, S, zero)
    assert _rel_err(pos.grad, gp_ref) < 1e-10
    assert _rel_err(cell.grad, gc_ref) < 1e-10

    pos2, cell2 = _setup(bcc)
    i, j, D, S = neighbour_list_autograd("ijDS", pos2, cell2, True, RC)
    w = torch.linspace(-1.0, 1.0, 3 * len(i), dtype=torch.float64,
                       device="cuda:0").reshape(-1, 3)
    (w * D).sum().backward()
    # dL/dr[k] = sum over pairs with j == k of w - sum over pairs with i == k of w
    ref = torch.zeros_like(pos2)
    ref.index_add_(0, j.long(), w)
    ref.index_add_(0, i.long(), -w)
    assert _rel_err(pos2.grad, ref) < 1e-12
    gc = S.to(torch.float64).T @ w
    assert _rel_err(cell2.grad, gc) < 1e-12


def _local_loss(pos_np, cell_np, k):
    """sum (1 - d/rc)^3 over the pairs that involve atom k (the only terms
    whose derivative with respect to r[k] is non-zero), from a fresh build."""
    from matscipy_neighbours_amd import neighbour_list
    i, j, d = neighbour_list("ijd", positions=pos_np, cell=cell_np, pbc=True, cutoff=RC)
    m = (i == k) | (j == k)
    return float(((1.0 - d[m] / RC) ** 3).sum())


def _full_loss(pos_np, cell_np):
    from matscipy_neighbours_amd import neighbour_list
    d = neighbour_list("d", positions=pos_np, cell=cell_np, pbc=True, cutoff=RC)
    return float(np.sum((1.0 - d / RC) ** 3))


def test_finite_differences(gpu, bcc):
    """Central differences, h = 1e-6 A, float64, rebuilding the list."""
    from matscipy_neighbours_amd.autograd import neighbour_list_autograd
    pos, cell = _setup(bcc)
    d = neighbour_list_autograd("d", pos, cell, True, RC)
    _loss(d).backward()
    gp = pos.grad.cpu().numpy()
    gc = cell.grad.cpu().numpy()
    h = 1e-6
    P0, C0 = bcc["positions"], bcc["cell"]
    rng = np.random.default_rng(17)
    for k in rng.choice(len(P0), 6, replace=False):
        for a in range(3):
            pp, pm = P0.copy(), P0.copy()
            pp[k, a] += h
            pm[k, a] -= h
            fd = (_local_loss(pp, C0, k) - _local_loss(pm, C0, k)) / (2 * h)
            assert abs(fd - gp[k, a]) < 1e-6 * max(1.0, abs(gp[k, a])), (k, a, fd, gp[k, a])
    for a, b in ((0, 0), (1, 0), (1, 1), (2, 2), (0, 2)):
        cp, cm = C0.copy(), C0.copy()
        cp[a, b] += h
        cm[a, b] -= h
        fd = (_full_loss(P0, cp) - _full_loss(P0, cm)) / (2 * h)
        # the full loss sums ~1e7 terms of size ~0.1: its rounding error
        # (~1e-10) over 2h bounds the difference at ~1e-4
        assert abs(fd - gc[a, b]) < 1e-3 + 1e-6 * abs(gc[a, b]), (a, b, fd, gc[a, b])


def test_float32_forward_and_backward(gpu, bcc):
    import torch

    from matscipy_neighbours_amd.autograd import neighbour_list_autograd
    pos, cell = _setup(bcc)
    i, j, D, d, S = neighbour_list_autograd("ijDdS", pos, cell, True, RC,
                                            out_dtype=torch.float32)
    assert D.dtype == torch.float32 and d.dtype == torch.float32
    pos64, cell64 = _setup(bcc)
    i64, j64, D64, d64, S64 = neighbour_list_autograd("ijDdS", pos64, cell64, True, RC)
    # same pair set, D and d within 2e-6 relative of the float64 values
    assert torch.equal(i, i64) and torch.equal(j, j64) and torch.equal(S, S64)
    Dd, D64d = D.detach().double(), D64.detach()
    assert float(((Dd - D64d).abs() / D64d.abs().clamp_min(1e-30)).max()) <= 2e-6 or \
        float((Dd - D64d).abs().max()) <= 2e-6 * float(D64d.abs().max())
    assert float(((d.detach().double() - d64.detach()) / d64.detach()).abs().max()) <= 2e-6
    _loss(d.double()).backward()
    _loss(d64).backward()
    assert pos.grad.dtype == torch.float64
    assert _rel_err(pos.grad, pos64.grad) < 1e-5
    assert _rel_err(cell.grad, cell64.grad) < 1e-5


def test_gradcheck_small_triclinic(gpu):
    import torch

    from matscipy_neighbours_amd.autograd import neighbour_list_autograd
    rng = np.random.default_rng(4)
    cellnp = np.array([[5.0, 0.0, 0.0], [0.9, 4.6, 0.0], [0.4, -0.7, 5.3]])
    posnp = rng.uniform(0, 1, (24, 3)) @ cellnp
    pos = torch.tensor(posnp, dtype=torch.float64, device="cuda:0", requires_grad=True)
    cell = torch.tensor(cellnp, dtype=torch.float64, device="cuda:0", requires_grad=True)

    def f(p, c):
        D, d = neighbour_list_autograd("Dd", p, c, (True, True, False), 2.7)
        return D, d

    assert torch.autograd.gradcheck(f, (pos, cell), eps=1e-6, atol=1e-7, rtol=1e-6)


def test_cell_gradient_ignores_nonperiodic_shift(gpu):
    """pbc (T, T, F) with atoms more than a box length outside along z: the
    reference's S then holds the home atom's image count there (its
    truncating (raw_i - rel_j + x)/n, neighbour_visit.hh:111-114), while D
    is r_j - r_i along z. The cell gradient follows D, i.e. torch autograd
    of r[j] - r[i] + (S * pbc) @ cell."""
    import torch

    from matscipy_neighbours_amd.autograd import neighbour_list_autograd
    rng = np.random.default_rng(21)
    cellnp = np.array([[6.0, 0.0, 0.0], [0.8, 5.7, 0.0], [0.3, -0.4, 6.2]])
    frac = rng.uniform(0, 1, (60, 3))
    frac[:20, 2] = rng.uniform(-2.6, -1.2, 20)   # one to two and a half boxes below
    frac[20:40, 2] = rng.uniform(2.2, 3.5, 20)   # above
    posnp = frac @ cellnp
    pbc = (True, True, False)
    pos = torch.tensor(posnp, dtype=torch.float64, device="cuda:0", requires_grad=True)
    cell = torch.tensor(cellnp, dtype=torch.float64, device="cuda:0", requires_grad=True)
    i, j, D, d, S = neighbour_list_autograd("ijDdS", pos, cell, pbc, 3.0)
    assert int((S[:, 2] != 0).sum()) > 0, "no pair carries a non-periodic image count"
    w = torch.randn(D.shape, dtype=torch.float64, device="cuda:0",
                    generator=torch.Generator(device="cuda:0").manual_seed(5))
    L = _loss(d) + (w * D).sum()
    L.backward()
    mask = torch.tensor([1.0, 1.0, 0.0], dtype=torch.float64, device="cuda:0")
    p = pos.detach().clone().requires_grad_(True)
    c = cell.detach().clone().requires_grad_(True)
    Dr = p[j.long()] - p[i.long()] + (S.to(torch.float64) * mask) @ c
    assert _rel_err(D.detach(), Dr.detach()) < 1e-12
    (_loss(torch.linalg.norm(Dr, dim=1)) + (w * Dr).sum()).backward()
    assert _rel_err(pos.grad, p.grad) < 1e-10
    assert _rel_err(cell.grad, c.grad) < 1e-10
