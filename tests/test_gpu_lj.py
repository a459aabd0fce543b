"""The fused Lennard-Jones consumer (forces, energy and virial from the cell
grid, no pair list) against the same sums over the oracle's pair list.

For every pair of the list (i, j, D, d): s6 = (sigma^2 / d^2)^3,
coef = -24 eps (2 s6^2 - s6) / d^2, f_i += coef D (the force loop of the
reference's examples/lj_langevin/lj_langevin_cpu.cc:76-90, with eps and
sigma), E = 1/2 sum 4 eps (s6^2 - s6), W = -1/2 sum coef D (x) D. The
engine adds each atom's terms in another order (float64 atomics), so the
forces agree to a tolerance scaled by each atom's sum of absolute terms,
written below; energy and virial likewise."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-11


def _from_list(i, D, d, n, eps, sig):
    r2 = d * d
    s6 = (sig * sig / r2) ** 3
    coef = -24.0 * eps * (2.0 * s6 * s6 - s6) / r2
    terms = coef[:, None] * D
    f = np.zeros((n, 3))
    np.add.at(f, i, terms)
    fabs = np.zeros(n)
    np.add.at(fabs, i, np.abs(terms).sum(1))
    e_terms = 0.5 * 4.0 * eps * (s6 * s6 - s6)
    w_terms = -0.5 * coef[:, None, None] * D[:, :, None] * D[:, None, :]
    return (f, fabs, e_terms.sum(), np.abs(e_terms).sum(), w_terms.sum(0),
            np.abs(w_terms).sum(0))


def _compare(pos, cell, pbc, rc, eps, sig):
    from matscipy_neighbours_amd import lennard_jones
    inv = np.ascontiguousarray(np.linalg.inv(np.asarray(cell, float).T))
    ref = oracle.neighbour_list("ijdD", np.zeros(3), np.asarray(cell, float), inv, pbc,
                                np.ascontiguousarray(pos), rc)
    f0, fabs, e0, eabs, w0, wabs = _from_list(ref["i"], ref["D"], ref["d"], len(pos), eps, sig)
    e, f, w = lennard_jones(positions=pos, cell=cell, pbc=pbc, cutoff=rc, epsilon=eps,
                            sigma=sig)
    assert f.shape == (len(pos), 3) and w.shape == (3, 3)
    assert np.all(np.abs(f - f0) <= TOL * (fabs[:, None] + 1e-300))
    assert abs(e - e0) <= TOL * eabs
    assert np.all(np.abs(w - w0) <= TOL * wabs + 1e-300)
    assert np.array_equal(w, w.T)
    return e, f, w


@pytest.mark.parametrize("pbc", [True, (True, False, True)])
def test_lj_random_periodic(gpu, pbc):
    rng = np.random.default_rng(81)
    cell = np.diag([9.0, 10.0, 11.0])
    pos = rng.uniform(0.0, 1.0, (600, 3)) @ cell
    _compare(pos, cell, pbc, 2.5, 0.8, 1.1)


def test_lj_triclinic_small_box(gpu):
    """A box below twice the cutoff: pairs with several periodic images."""
    rng = np.random.default_rng(82)
    cell = np.array([[4.0, 0.0, 0.0], [1.0, 4.5, 0.0], [0.5, 0.7, 5.0]])
    pos = rng.uniform(0.0, 1.0, (40, 3)) @ cell
    _compare(pos, cell, True, 3.0, 1.0, 1.0)


def test_lj_droplet_open_boundaries(gpu):
    """The reference example's setting: an fcc droplet (lattice 1.6) in an
    open box, eps = sigma = 1, cutoff 2.5."""
    a, k = 1.6, 6
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(*[np.arange(k)] * 3, indexing="ij"), -1).reshape(-1, 1, 3)
    pos = ((g + basis[None]) * a).reshape(-1, 3)
    centre = pos.mean(0)
    pos = pos[np.linalg.norm(pos - centre, axis=1) < k * a / 2]
    pos = pos + np.random.default_rng(83).normal(0.0, 0.02, pos.shape)
    box = np.diag(pos.max(0) - pos.min(0) + 6.0)
    pos = pos - pos.min(0) + 3.0
    e, f, w = _compare(pos, box, False, 2.5, 1.0, 1.0)
    # open boundaries, no external field: the forces cancel
    assert np.abs(f.sum(0)).max() <= 1e-9 * np.abs(f).sum()
    assert e < 0.0


def test_lj_full_size_against_engine_list(gpu):
    """fcc 1M (a = 3.615, rc 5 A) with sigma 2.3: the fused sums against the
    same sums over the engine's own ijdD list on the device (the list is
    bit-exact against the oracle elsewhere), and the forces cancel."""
    import torch

    from matscipy_neighbours_amd import lennard_jones, neighbour_list, workloads
    w = workloads.fcc_bulk()
    pos = torch.from_numpy(w["positions"]).to("cuda:0")
    cell = w["cell"]
    eps, sig = 0.4, 2.3
    e, f, vir = lennard_jones(positions=pos, cell=cell, pbc=True, cutoff=5.0, epsilon=eps,
                              sigma=sig)
    i, D, d = neighbour_list("iDd", positions=pos, cell=cell, pbc=True, cutoff=5.0)
    i = torch.from_dlpack(i).long()
    D = torch.from_dlpack(D)
    d = torch.from_dlpack(d)
    r2 = d * d
    s6 = (sig * sig / r2) ** 3
    coef = -24.0 * eps * (2.0 * s6 * s6 - s6) / r2
    terms = coef[:, None] * D
    f0 = torch.zeros_like(f).index_add_(0, i, terms)
    fabs = torch.zeros(f.shape[0], dtype=torch.float64, device=f.device).index_add_(
        0, i, terms.abs().sum(1))
    assert bool((torch.abs(f - f0) <= TOL * fabs[:, None]).all())
    e0 = (0.5 * 4.0 * eps * (s6 * s6 - s6)).sum()
    assert abs(float(e) - float(e0)) <= TOL * float((2.0 * eps * (s6 * s6 + s6)).sum())
    w0 = (-0.5 * coef[:, None, None] * D[:, :, None] * D[:, None, :]).sum(0)
    wabs = (0.5 * coef.abs()[:, None, None] * (D[:, :, None] * D[:, None, :]).abs()).sum(0)
    assert bool((torch.abs(vir - w0) <= TOL * wabs).all())
    assert float(f.sum(0).abs().max()) <= 1e-9 * float(f.abs().sum())


def test_lj_edges(gpu):
    from matscipy_neighbours_amd import lennard_jones
    e, f, w = lennard_jones(positions=np.zeros((0, 3)), cell=np.eye(3) * 5, pbc=True,
                            cutoff=2.0)
    assert e == 0.0 and f.shape == (0, 3) and np.all(w == 0.0)
    with pytest.raises(TypeError):
        lennard_jones(positions=np.zeros((2, 3)), cell=np.eye(3) * 5, pbc=True,
                      cutoff=np.array([1.0, 1.0]))


def _lattice(spacing=1.2, k=6, jitter=0.03, seed=84):
    g = np.stack(np.meshgrid(*[np.arange(k)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = g * spacing + np.random.default_rng(seed).normal(0.0, jitter, (len(g), 3))
    return pos, np.eye(3) * k * spacing


def test_lj_dimer_known_answers(gpu):
    """Two atoms in an open box: at r = 2^(1/6) sigma the energy is -eps and
    the force vanishes; at r = sigma the energy is 0 and the force is
    24 eps / sigma, pushing the atoms apart."""
    from matscipy_neighbours_amd import lennard_jones
    eps, sig = 0.7, 1.3
    cell = np.eye(3) * 20.0
    for r, e_want, f_want in [(2 ** (1 / 6) * sig, -eps, 0.0), (sig, 0.0, 24 * eps / sig)]:
        pos = np.array([[5.0, 6.0, 7.0], [5.0, 6.0, 7.0]]) + np.array([[0, 0, 0], [r, 0, 0]])
        e, f, w = lennard_jones(positions=pos, cell=cell, pbc=False, cutoff=3 * sig,
                                epsilon=eps, sigma=sig)
        assert abs(e - e_want) <= 1e-13
        assert np.allclose(f[0], [-f_want, 0, 0], rtol=0, atol=1e-12)
        assert np.allclose(f[1], [f_want, 0, 0], rtol=0, atol=1e-12)
        # one pair along x: W = r_01 (x) f_01 = (-r) * (-f_want) in xx only
        assert abs(w[0, 0] - r * f_want) <= 1e-12 and np.all(np.delete(w.ravel(), 0) == 0)


def test_lj_forces_are_minus_energy_gradient(gpu):
    """f = -dE/dr by central differences (h = 1e-5) on a jittered simple
    cubic lattice (spacing 1.2, 0.03 jitter, rc 2.2), periodic. With these
    seeds no pair lies within 1.6e-3 of the cutoff (0.0016 for seed 84,
    0.0034 for seed 85, measured), so no pair crosses it under the
    perturbations here."""
    from matscipy_neighbours_amd import lennard_jones
    pos, cell = _lattice()
    kw = dict(cell=cell, pbc=True, cutoff=2.2, epsilon=1.0, sigma=1.0)
    e, f, _ = lennard_jones(positions=pos, **kw)
    h = 1e-5
    for a in (0, 17, 101, 215):
        for k in range(3):
            p, m = pos.copy(), pos.copy()
            p[a, k] += h
            m[a, k] -= h
            fd = -(lennard_jones(positions=p, **kw)[0] - lennard_jones(positions=m, **kw)[0]) / (2 * h)
            assert abs(fd - f[a, k]) <= 1e-6 * (1.0 + abs(f[a, k])), (a, k, fd, f[a, k])


def test_lj_virial_is_minus_strain_derivative(gpu):
    """W = -dE/d(strain) for a homogeneous symmetric strain of positions and
    cell together, by central differences (h = 1e-6: displacements below
    1e-5, far inside the lattice's distance to the cutoff)."""
    from matscipy_neighbours_amd import lennard_jones
    pos, cell = _lattice(seed=85)
    kw = dict(pbc=True, cutoff=2.2, epsilon=1.0, sigma=1.0)
    _, _, w = lennard_jones(positions=pos, cell=cell, **kw)
    h = 1e-6
    for a, b in [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]:
        eps_ab = np.zeros((3, 3))
        eps_ab[a, b] = eps_ab[b, a] = 0.5 if a != b else 1.0
        es = []
        for sgn in (1, -1):
            F = np.eye(3) + sgn * h * eps_ab
            es.append(lennard_jones(positions=pos @ F.T, cell=cell @ F.T, **kw)[0])
        # d/dh E((I + h e) r) = sum_cd dE/d(strain_cd) e_cd = dE/d(strain_ab),
        # the two halves of a symmetric off-diagonal e adding up
        want = -(es[0] - es[1]) / (2 * h)
        assert abs(w[a, b] - want) <= 1e-5 * (1.0 + abs(w[a, b])), (a, b, w[a, b], want)


def test_lj_sparse_box_hashed_grid(gpu):
    """A sparse cloud in a 600 A box at rc 3 (200^3 cells, more than the dense
    budget allows): the build takes the hashed grid, and the fused sums still
    match the oracle list's."""
    from matscipy_neighbours_amd import _lib
    rng = np.random.default_rng(86)
    cell = np.eye(3) * 600.0
    centres = rng.uniform(50.0, 550.0, (30, 3))
    pos = (centres[:, None, :] + rng.normal(0.0, 2.5, (30, 40, 3))).reshape(-1, 3)
    _compare(pos, cell, False, 3.0, 0.9, 1.2)
    plan = _lib.plan_for(0)
    assert _lib.load().mnl_plan_backend(plan.handle) == 1
