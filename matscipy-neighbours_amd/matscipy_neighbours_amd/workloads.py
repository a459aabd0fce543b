"""Seeded synthetic inputs for the neighbour-list workloads.

Each generator returns a dict with positions (n, 3) float64, cell (3, 3) with
rows = lattice vectors, pbc (3,) bool, cutoff, and, where relevant, numbers.
"""

import numpy as np


def uniform_box(n=1000, edge=10.0, cutoff=2.5, seed=1, pbc=True):
    """n atoms uniform-random in a cubic periodic box (the 1000-atom,
    10 A, rc 2.5 A CPU-runnable case)."""
    rng = np.random.default_rng(seed)
    return dict(positions=rng.uniform(0.0, edge, (n, 3)), cell=edge * np.eye(3),
                pbc=np.array([pbc] * 3) if np.isscalar(pbc) else np.asarray(pbc, bool),
                cutoff=float(cutoff))


def fcc_bulk(reps=64, a=3.615, sigma=0.1, cutoff=5.0, seed=2):
    """fcc lattice, reps^3 cubic unit cells (4 sites each, C order), Gaussian
    jitter sigma per coordinate, orthorhombic periodic cell of edge reps*a.
    reps=64 gives 1,048,576 atoms in a 231.36 A cell (~46M pairs at rc 5)."""
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.arange(reps, dtype=np.float64)
    cells = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 1, 3)
    pos = ((cells + basis[None, :, :]) * a).reshape(-1, 3)
    rng = np.random.default_rng(seed)
    pos = pos + rng.normal(0.0, sigma, pos.shape)
    return dict(positions=np.ascontiguousarray(pos), cell=reps * a * np.eye(3),
                pbc=np.array([True, True, True]), cutoff=float(cutoff))


def triclinic_slab(n=200_000, seed=3):
    """Two species (random 50/50) uniform in the bottom quarter of the
    triclinic cell a=(120,0,0), b=(12,118,0), c=(5,6,400), vacuum above,
    pbc (T,T,F), per-type cutoff matrix [[3.0,3.5],[3.5,4.0]]."""
    rng = np.random.default_rng(seed)
    cell = np.array([[120.0, 0, 0], [12.0, 118.0, 0], [5.0, 6.0, 400.0]])
    frac = np.column_stack([rng.uniform(0, 1, n), rng.uniform(0, 1, n),
                            rng.uniform(0, 0.25, n)])
    numbers = rng.integers(0, 2, n).astype(np.int32)
    return dict(positions=np.ascontiguousarray(frac @ cell), cell=cell,
                pbc=np.array([True, True, False]), numbers=numbers,
                cutoff=np.array([[3.0, 3.5], [3.5, 4.0]]))


def vacuum_cluster(n=131_072, density=0.08, factor=20.0, cutoff=5.0, seed=6):
    """n atoms uniform in a cube of edge (n/density)^(1/3) in the corner of a
    periodic cubic box `factor` times wider: mostly vacuum, so the cell grid
    (about 471^3 cells at the defaults) exceeds max(2^20, 8n) and the
    reference's rule picks the hashed grid (the shape of the reference's
    sparse-vacuum test, tests/test_neighbour_list_gpu.cc:116-123, at a larger n)."""
    rng = np.random.default_rng(seed)
    edge = (n / density) ** (1.0 / 3.0)
    return dict(positions=rng.uniform(0.0, edge, (n, 3)), cell=factor * edge * np.eye(3),
                pbc=np.array([True, True, True]), cutoff=float(cutoff))


def bcc_sheared(reps=(64, 32, 32), a=2.87, sigma=0.08, shear=0.05, cutoff=6.0, seed=5):
    """bcc lattice (2 sites per cubic cell), reps cells, Gaussian jitter,
    periodic cell diag(reps*a) with an xy shear; 64x32x32 gives 131,072 atoms."""
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0.5]])
    gx, gy, gz = (np.arange(r, dtype=np.float64) for r in reps)
    cells = np.stack(np.meshgrid(gx, gy, gz, indexing="ij"), axis=-1).reshape(-1, 1, 3)
    frac_lat = ((cells + basis[None]) / np.array(reps, dtype=float)).reshape(-1, 3)
    cell = np.diag(np.array(reps, dtype=float) * a)
    cell[1, 0] = shear * cell[1, 1]
    rng = np.random.default_rng(seed)
    pos = frac_lat @ cell + rng.normal(0.0, sigma, (frac_lat.shape[0], 3))
    return dict(positions=np.ascontiguousarray(pos), cell=cell,
                pbc=np.array([True, True, True]), cutoff=float(cutoff))


def uneven_batch(nsys=1024, nmin=8, nmax=512, density=0.05, tilt=0.1, cutoff=5.0,
                 seed=4):
    """nsys periodic systems with atom counts uniform in [nmin, nmax], cubic
    cells sized for `density` atoms/A^3 with random off-diagonal tilt up to
    `tilt` of the edge, uniform fractional positions."""
    rng = np.random.default_rng(seed)
    systems = []
    for _ in range(nsys):
        n = int(rng.integers(nmin, nmax + 1))
        edge = (n / density) ** (1.0 / 3.0)
        t = rng.uniform(-tilt, tilt, (3, 3))
        np.fill_diagonal(t, 0.0)
        cell = edge * (np.eye(3) + t)
        frac = rng.uniform(0.0, 1.0, (n, 3))
        systems.append(dict(positions=np.ascontiguousarray(frac @ cell), cell=cell,
                            pbc=np.array([True, True, True]), cutoff=float(cutoff)))
    return systems
