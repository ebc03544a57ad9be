"""ACA (src/ACA.cpp:127-238) as restated by the oracle: invariants, known
answers and the effect of the reduction order (CPU)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import lattice, uniform


def dense(xyz_pos, rs, n, cs, m, q=7):
    a = xyz_pos[rs:rs + n][:, None, :]
    b = xyz_pos[cs:cs + m][None, :, :]
    d = a - b
    r = np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2])
    if q == 7:
        return 1 / r
    if q == 3:
        return 1 + r * r
    if q == 1:
        return np.exp(-r)
    raise ValueError(q)


def setup(pts, leaf1d=4, q=7, tol=8):
    o = O.Oracle(pts, leaf1d, tol_pow=tol, qchoice=q, keep_uv=True)
    P = pts[o.perm()]
    return o, P


def blocks(o, level, count, seed=0):
    s, c, _ = o.boxes(level)
    off, idx = o.il(level)
    rng = np.random.default_rng(seed)
    out = []
    for k in rng.choice(8 ** level, count, replace=False):
        for b in range(off[k], off[k + 1], 7):
            ki = idx[b]
            if c[k] and c[ki]:
                out.append((s[k], c[k], s[ki], c[ki]))
    return out


def test_cross_interpolation_and_factor_identities():
    o, P = setup(uniform(4096, 1))
    for rs, n, cs, m in blocks(o, 2, 6):
        r, I, J, U, V = o.aca_positions(rs, n, cs, m)
        A = dense(P, rs, n, cs, m)
        Ap = U.T @ V
        # the skeleton rows and columns are reproduced (to roundoff)
        scale = np.abs(A).max()
        assert np.abs(Ap[I, :] - A[I, :]).max() < 1e-12 * scale
        assert np.abs(Ap[:, J] - A[:, J]).max() < 1e-12 * scale
        # v_l vanishes at earlier pivot columns; u_l at earlier pivot rows
        for l in range(r):
            assert np.all(V[l, J[:l]] == 0.0) or np.abs(V[l, J[:l]]).max() < 1e-13 * scale
        # pivots are distinct
        assert len(set(I)) == r and len(set(J)) == r
        # accuracy at the requested tolerance (weak-admissible blocks)
        assert np.linalg.norm(A - Ap) <= 50 * 1e-8 * np.linalg.norm(A)


def test_known_answers():
    o, P = setup(uniform(2048, 2), q=3, tol=12)
    for rs, n, cs, m in blocks(o, 2, 4):
        r, I, J, U, V = o.aca_positions(rs, n, cs, m)
        A = dense(P, rs, n, cs, m, q=3)
        # 1 + r^2 = 1 + |x|^2 + |y|^2 - 2 x.y has rank <= 5
        assert r <= 6
        assert np.linalg.norm(A - U.T @ V) < 1e-12 * np.linalg.norm(A)


def test_full_rank_reproduces_block():
    # tiny boxes (one point per leaf -> rank 1 == min(n, m)); 8-point boxes at level 1
    pts = uniform(64, 3)
    o = O.Oracle(pts, 1, tol_pow=16, keep_uv=True, nlevels=2)
    P = pts[o.perm()]
    s, c, _ = o.boxes(1)
    r, I, J, U, V = o.aca_positions(s[0], c[0], s[6], c[6])
    A = dense(P, s[0], c[0], s[6], c[6])
    assert r <= min(c[0], c[6])
    assert np.linalg.norm(A - U.T @ V) < 1e-13 * np.linalg.norm(A)


def test_zero_block_has_rank_zero():
    pts = uniform(512, 4)
    o = O.Oracle(pts, 2, qchoice=7, scale=0.0)
    o.assemble()
    for j in range(1, o.nLevels + 1):
        r, st = o.ranks(j)
        assert r.max() == 0


def test_empty_column_box_gives_rank_zero():
    # all points in one octant: most boxes are empty (UB in the reference ACA)
    pts = uniform(3000, 5, 0.05, 0.95)
    o = O.Oracle(pts, 4)
    o.assemble()
    y = o.matvec(np.ones(len(pts)))
    assert np.all(np.isfinite(y))


def test_first_pivot_and_tie_rules():
    # lattice blocks have many exact ties in |K|; the first maximum must win
    pts = lattice(8)
    o = O.Oracle(pts, 2, tol_pow=10, keep_uv=True)
    s, c, _ = o.boxes(1)
    r, I, J, U, V = o.aca_positions(s[0], c[0], s[6], c[6])
    P = pts[o.perm()]
    A = dense(P, s[0], c[0], s[6], c[6])
    row0 = np.abs(A[0])
    assert I[0] == 0 and J[0] == int(np.argmax(row0))  # np.argmax returns the first maximum
    col0 = np.abs(A[:, J[0]]).copy()
    col0[I[0]] = -1
    assert I[1] == int(np.argmax(col0))


def test_reduction_order_rarely_changes_decisions():
    """The canonical 256-lane order vs Eigen's 2-wide packet order: pivots are
    unaffected (argmax does not reduce) and stop decisions differ at most in
    near-threshold cases."""
    pts = lattice(16)
    ranks = []
    for mode in (0, 1):
        O.set_reduction(mode)
        try:
            o = O.Oracle(pts, 4, tol_pow=10)
            o.assemble()
            ranks.append(np.concatenate([o.ranks(j)[0] for j in (1, 2)]))
        finally:
            O.set_reduction(0)
    diff = int((ranks[0] != ranks[1]).sum())
    assert diff <= len(ranks[0]) // 1000
