"""Input generators shared by the tests, the smoke entry and the benchmark.

They restate the snapshot factories of the reference test-suite (hoomd/conftest.py:119-304)
and the workloads of the benchmark configs; they hold no expected outputs.
"""

import numpy as np


def sc_lattice(n=7, a=1.0):
    """lattice_snapshot_factory (hoomd/conftest.py:202-262), 3D, no perturbation.
    Returns (positions[N,3], L)."""
    ranges = [np.arange(-n / 2, n / 2) for _ in range(3)]
    x, y, z = np.meshgrid(*ranges)
    pos = (np.vstack((x.flatten(), y.flatten(), z.flatten())).T + 0.5) * a
    return pos, n * a


def fcc_lattice(n, rho=0.8442):
    """FCC n^3 x 4 at number density rho in the cubic box [-L/2, L/2)^3 (4 basis sites
    (0,0,0), (0,a/2,a/2), (a/2,0,a/2), (a/2,a/2,0) per unit cell; hoomd/conftest.py:265-304).
    Returns (positions[N,3], L)."""
    a = (4.0 / rho) ** (1.0 / 3.0)
    L = n * a
    basis = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]]) * a
    g = np.arange(n) * a
    cx, cy, cz = np.meshgrid(g, g, g, indexing="ij")
    cells = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1)
    pos = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3) - L / 2 + a / 4
    return pos, L


def random_fluid(N, rho, seed=1, min_dist=None):
    """Uniform random positions in a cubic box at density rho (for list-building tests)."""
    L = (N / rho) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-L / 2, L / 2, size=(N, 3))
    return pos, L


def perturbed_fcc(n, rho=0.8442, r=0.05, seed=3):
    pos, L = fcc_lattice(n, rho)
    rng = np.random.default_rng(seed)
    pos = pos + rng.uniform(-r, r, size=pos.shape)
    pos = (pos + L / 2) % L - L / 2
    return pos, L


def maxwell_boltzmann(N, kT, seed=7, mass=None):
    """Gaussian velocities with zero COM momentum (numpy draws; the Philox-based
    thermalize of the reference is exercised separately)."""
    rng = np.random.default_rng(seed)
    m = np.ones(N) if mass is None else np.asarray(mass, np.float64)
    v = rng.normal(0.0, 1.0, size=(N, 3)) * np.sqrt(kT / m)[:, None]
    p = (m[:, None] * v).sum(axis=0) / N
    return v - p[None, :] / m[:, None]


def random_walk_chains(n_chains, chain_len, rho, bond=0.97, seed=1):
    """Linear chains as random walks with fixed bond length in a cubic box at bead density rho
    (the Kremer-Grest melt's initial state: ideal chains that overlap freely). Returns
    (positions[N,3] wrapped into [-L/2, L/2), images[N,3], bonds[M,2] by tag, L); tags run
    chain by chain."""
    N = n_chains * chain_len
    L = (N / rho) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    start = rng.uniform(-L / 2, L / 2, size=(n_chains, 1, 3))
    v = rng.normal(size=(n_chains, chain_len - 1, 3))
    v *= bond / np.linalg.norm(v, axis=2, keepdims=True)
    steps = np.concatenate([np.zeros((n_chains, 1, 3)), np.cumsum(v, axis=1)], axis=1)
    unwrapped = (start + steps).reshape(N, 3)
    image = np.floor((unwrapped + L / 2) / L).astype(np.int32)
    pos = unwrapped - image * L
    up = pos >= L / 2
    pos[up] -= L
    image[up] += 1
    first = (np.arange(n_chains)[:, None] * chain_len + np.arange(chain_len - 1)[None, :]).ravel()
    bonds = np.stack([first, first + 1], axis=1).astype(np.uint32)
    return pos, image, bonds, L
