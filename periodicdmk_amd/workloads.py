"""Synthetic inputs for the benchmark and parity workloads.

Random numbers come from splitmix64 with u = (x >> 11) * 2^-53, so the same seed
gives the same positions and charges in Python and C++ on any machine.

Workloads (cubic cells have L = 1):
  * ``uniform_cubic(n, seed)``            -- positions i.i.d. uniform in the cell;
  * ``neutral_uniform_charges(n, seed)``  -- i.i.d. U[-1,1], mean subtracted;
  * ``gaussian_mixture(...)``             -- isotropic Gaussian clusters, wrapped
    periodically in FP64 fractional coordinates;
  * ``pm1_charges(n, seed)``              -- +/-1 with equal counts, shuffled;
  * ``batch_cells(...)``                  -- the batched workload's cells and point counts.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """The first ``n`` outputs of splitmix64 started from ``seed`` (uint64)."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int) -> np.ndarray:
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def cubic_cell(L: float = 1.0) -> np.ndarray:
    """Cell matrix with the lattice vectors as columns."""
    return np.eye(3) * L


def triclinic_cell() -> np.ndarray:
    """Columns a1=(1,0,0), a2=(0.3,1,0), a3=(0.2,0.25,1)."""
    return np.array([[1.0, 0.3, 0.2], [0.0, 1.0, 0.25], [0.0, 0.0, 1.0]])


def uniform_positions(n: int, seed: int, cell=None) -> np.ndarray:
    """(n, 3) Cartesian positions, uniform in fractional coordinates of ``cell``."""
    cell = cubic_cell() if cell is None else np.asarray(cell, dtype=np.float64)
    frac = uniform01(seed, 3 * n).reshape(n, 3)
    return frac @ cell.T


def neutral_uniform_charges(n: int, seed: int) -> np.ndarray:
    q = 2.0 * uniform01(seed, n) - 1.0
    return q - q.mean()


def pm1_charges(n: int, seed: int) -> np.ndarray:
    q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    perm = np.argsort(uniform01(seed, n), kind="stable")
    return q[perm]


def gaussian_mixture(n: int, n_clusters: int, sigma: float, seed: int, cell=None) -> np.ndarray:
    """Points from ``n_clusters`` isotropic Gaussians (std ``sigma`` in Cartesian units),
    centres uniform in fractional coordinates, wrapped into the cell in FP64."""
    cell = cubic_cell() if cell is None else np.asarray(cell, dtype=np.float64)
    u = uniform01(seed, 3 * n_clusters + n + 6 * n)
    centres = u[: 3 * n_clusters].reshape(n_clusters, 3) @ cell.T
    which = np.minimum((u[3 * n_clusters: 3 * n_clusters + n] * n_clusters).astype(np.int64), n_clusters - 1)
    g = u[3 * n_clusters + n:].reshape(n, 6)
    # Box-Muller on pairs
    r = np.sqrt(-2.0 * np.log(np.maximum(g[:, 0:3], 2.0 ** -53)))
    z = r * np.cos(2.0 * np.pi * g[:, 3:6])
    x = centres[which] + sigma * z
    frac = np.linalg.solve(cell, x.T).T
    frac -= np.floor(frac)
    frac[frac >= 1.0] = 0.0
    return frac @ cell.T


def batch_cells(n_cells: int = 256, seed: int = 31, n_min: int = 2_000, n_max: int = 200_000):
    """Cells and point counts of the batched workload (many small trees): cell i is cubic with
    L = 1 when u0 < 1/2, else triclinic with columns (1,0,0), (t1,1,0), (t2,t3,1), t ~ U[0, 0.3];
    N_i is log-uniform in [n_min, n_max].  Points of cell i: ``uniform_positions(N_i, 1000 + i,
    cell)``; charge vector v of cell i: ``neutral_uniform_charges(N_i, 100000 + 32 i + v)``."""
    u = uniform01(seed, 5 * n_cells).reshape(n_cells, 5)
    out = []
    for i in range(n_cells):
        cell = cubic_cell()
        if u[i, 0] >= 0.5:
            cell[0, 1], cell[0, 2], cell[1, 2] = 0.3 * u[i, 1], 0.3 * u[i, 2], 0.3 * u[i, 3]
        n = int(round(np.exp(np.log(n_min) + u[i, 4] * (np.log(n_max) - np.log(n_min)))))
        out.append((cell, n))
    return out
