"""Synthetic point sets and operators of the benchmark configurations.

All generators are seeded (numpy PCG64) so the GPU engine, the CPU oracle and
the direct sums see identical FP64 inputs.
  lattice(n)      the reference's own generator (examples/testHODLR3D.cpp:15-31)
  uniform(N)      iid uniform in [-1,1]^3
  jittered(N)     N distinct cells of the cell-centred 128^3 lattice of
                  [-1,1]^3 (spacing h = 2/128), each point jittered by
                  U(+-0.25 h): keeps points off every dyadic split plane
"""
import math

import numpy as np


def lattice(n, L=1.0):
    t = -L + 2.0 * L * (np.arange(n) + 1.0) / (n + 1.0)
    j, k, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return np.stack([t[k.ravel()], t[j.ravel()], t[i.ravel()]], axis=1)


def uniform(N, seed=0, lo=-1.0, hi=1.0):
    return np.random.default_rng(seed).uniform(lo, hi, (N, 3))


def jittered(N, seed=0, n1d=128, jitter=0.25):
    rng = np.random.default_rng(seed)
    h = 2.0 / n1d
    cells = rng.choice(n1d ** 3, size=N, replace=False)
    cells.sort()
    iz, rem = np.divmod(cells, n1d * n1d)
    iy, ix = np.divmod(rem, n1d)
    base = np.stack([ix, iy, iz], axis=1).astype(np.float64)
    pts = -1.0 + (base + 0.5) * h + rng.uniform(-jitter * h, jitter * h, (N, 3))
    return pts


def second_kind_scale(n1d=128):
    """h^3/(4 pi) for A = I + h^3 K, K = 1/(4 pi r)."""
    h = 2.0 / n1d
    return h ** 3 / (4.0 * math.pi)
