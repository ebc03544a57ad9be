"""Kernel catalogue and quadrature of the oracle against library routines (CPU)."""
import math

import numpy as np
import pytest
from scipy import special

from oracle import oracle as O

R = [0.013, 0.1, 0.7, 1.3, 3.1]


def ref_phi(q, r):
    return {0: math.log(1 + r), 1: math.exp(-r), 2: 1 / (1 + r * r), 3: 1 + r * r,
            4: 1 / math.sqrt(r * r + 1), 5: math.exp(-r * r), 6: math.sqrt(r * r + 1), 7: 1 / r,
            8: 1 / r ** 4, 9: math.log(r), 10: r * r * math.log(r), 11: special.j0(r),
            12: special.y0(r), 13: math.cos(r) / r, 14: special.kv(0.5, r) / math.sqrt(r),
            15: math.exp(-50 * r) / r, 16: 1 / r}[q]


@pytest.mark.parametrize("q", range(17))
def test_catalogue(q):
    for r in R:
        a = np.array([0.1, -0.2, 0.3])
        b = a + np.array([r, 0, 0]) / math.sqrt(1.0) if q % 2 else a + r * np.array([0.6, 0.0, 0.8])
        v = O.entry(q, a, b)
        assert v == pytest.approx(ref_phi(q, r), rel=1e-12, abs=1e-300)


def test_diagonal_and_scale():
    a = np.array([0.1, 0.2, 0.3])
    assert O.entry(7, a, a, 5, 5) == 0.0  # i == j -> 0 (examples/testHODLR3D.cpp:34-36)
    assert O.entry(7, a, a, 5, 5, diag=2.5) == 2.5
    b = a + [0.5, 0, 0]
    assert O.entry(7, a, b, 0, 1, scale=3.0) == 3.0 * 2.0


def test_gauss_legendre_nodes():
    x, w = O.gauss_legendre(16)
    xr, wr = np.polynomial.legendre.leggauss(16)
    assert np.allclose(x, xr, atol=1e-15) and np.allclose(w, wr, atol=1e-15)


def test_ie_cube_self_term():
    # 1 + 8 * (16^3 Gauss-Legendre sum over [0,h/2]^3 of 1/(4 pi r)), h = N^(-1/3)
    N = 1000
    h = 1 / N ** (1 / 3)
    x, w = np.polynomial.legendre.leggauss(16)
    c = 0.25 * h
    t = c * x + c
    X, Y, Z = np.meshgrid(t, t, t, indexing="ij")
    W = w[:, None, None] * w[None, :, None] * w[None, None, :]
    val = 1 + 8 * c ** 3 * np.sum(W / (4 * np.pi * np.sqrt(X * X + Y * Y + Z * Z)))
    assert O.ie_cube_kii(N) == pytest.approx(val, rel=1e-13)
    # and it approaches 1 + the exact integral: int_{[0,a]^3} 1/r = a^2 * 1.190038...
    exact_unit = 3 * math.log((math.sqrt(3) + 1) / (math.sqrt(3) - 1)) / 2 - math.pi / 4
    a = h / 2
    assert (val - 1) == pytest.approx(8 * a * a * exact_unit / (4 * math.pi), rel=2e-2)


def test_plugin_entry_is_the_plummer_kernel():
    """Qchoice 100 is the plugin built in (include/h3d_plugin_plummer.hpp):
    1/sqrt(r^2 + eps^2), eps = 1/64, finite for coincident points."""
    eps = 1.0 / 64.0
    for r in R:
        a = np.array([0.1, -0.2, 0.3])
        b = a + r * np.array([0.6, 0.0, 0.8])
        d = a - b
        ref = 1.0 / math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + eps * eps)
        assert O.entry(100, a, b) == ref
    a = np.array([0.3, 0.3, 0.3])
    assert O.entry(100, a, a, 0, 1) == 64.0  # coincident points, distinct indices
    assert O.entry(100, a, a, 2, 2, diag=5.0) == 5.0  # i == j: the diagonal
