"""Oracle matvec and GMRES (src/HODLR3DTree.cpp:1647-1805, src/gmres.cpp) (CPU)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import lattice, uniform


@pytest.fixture(scope="module")
def c1():
    x = lattice(16)
    o = O.Oracle(x, 4, tol_pow=10, keep_uv=True)
    o.assemble()
    return x, o


def test_error_vs_direct_sum(c1):
    x, o = c1
    q = np.random.default_rng(0).uniform(-1, 1, len(x))
    y = o.matvec(q)
    yd = O.direct_rows(x, q, np.arange(len(x)))
    assert np.linalg.norm(y - yd) / np.linalg.norm(yd) < 1e-9


def test_direct_sum_matches_dense_numpy():
    x = uniform(700, 3)
    q = np.random.default_rng(1).standard_normal(700)
    d = x[:, None] - x[None]
    r = np.sqrt((d[..., 0] ** 2 + d[..., 1] ** 2) + d[..., 2] ** 2)
    np.fill_diagonal(r, np.inf)
    ref = (1 / r) @ q
    assert np.allclose(O.direct_rows(x, q, np.arange(700)), ref, rtol=1e-13, atol=0)


def test_cur_form_equals_aca_vectors(c1):
    x, o = c1
    q = np.random.default_rng(2).standard_normal(len(x))
    y0 = o.matvec(q, mode=0)
    y1 = o.matvec(q, mode=1)
    assert np.linalg.norm(y0 - y1) / np.linalg.norm(y0) < 1e-13


def test_linearity_and_columns(c1):
    x, o = c1
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(len(x)), rng.standard_normal(len(x))
    ya, yb, yab = o.matvec(a), o.matvec(b), o.matvec(2 * a - 3 * b)
    assert np.linalg.norm(yab - (2 * ya - 3 * yb)) / np.linalg.norm(yab) < 1e-13
    # column j: near rows are exact kernel entries (each pair is covered once)
    j = 1234
    e = np.zeros(len(x))
    e[j] = 1
    col = o.matvec(e, mode=1)
    d = x - x[j]
    r = np.sqrt((d[:, 0] ** 2 + d[:, 1] ** 2) + d[:, 2] ** 2)
    L = o.nLevels
    s, c, _ = o.boxes(L)
    perm = o.perm()
    pos = int(np.nonzero(perm == j)[0][0])
    lj = int(np.searchsorted(s, pos, side="right") - 1)
    nbr = set(o.neighbors(L)[lj]) | {lj}
    nbr.discard(-1)
    for k in nbr:
        rows = perm[s[k]:s[k] + c[k]]
        rows = rows[rows != j]
        assert np.array_equal(col[rows], 1 / r[rows])
    assert col[j] == 0.0


def test_reference_timers_and_memory(c1):
    x, o = c1
    a, m = o.timers()
    assert a > 0 and m > 0
    # findMemory: r*n + r^2 + m*r per block plus dense near field
    assert o.find_memory() > 0 and o.max_rank() > 0


def test_gmres_matches_dense_solve():
    x = lattice(10)
    N = len(x)
    kii = O.ie_cube_kii(N)
    o = O.Oracle(x, 2, tol_pow=12, qchoice=16, scale=1.0 / N, diag=kii)
    o.assemble()
    rng = np.random.default_rng(4)
    xs = rng.standard_normal(N)
    b = o.matvec(xs)
    xg, err, its, hist = o.gmres(b, 100, 1e-12)
    assert err <= 1e-12 and its < 100
    assert np.all(np.diff(hist) <= 1e-15)  # residual estimates are non-increasing
    assert np.linalg.norm(xg - xs) / np.linalg.norm(xs) < 1e-9
    # dense check of the same operator
    A = np.column_stack([o.matvec(np.eye(N)[:, i], mode=0) for i in range(0, N, 1)])
    xd = np.linalg.solve(A, b)
    assert np.linalg.norm(xg - xd) / np.linalg.norm(xd) < 1e-9
