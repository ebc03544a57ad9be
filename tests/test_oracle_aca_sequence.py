"""The oracle's ACA decision sequence pinned against dense blocks (CPU).

src/ACA.cpp:127-238 (LowRank::ACA_only_nodesCUR) is restated by the oracle;
these tests rebuild every decision of it from the dense block in numpy and the
returned vectors, so that a misreading shared by the GPU kernel and the oracle
(which compare bit for bit with each other) fails here:

* the residual row of step k is A[I_k,:] - sum_{l<k} u_l[I_k] v_l, updated with
  l outermost and no fused multiply-add (ACA.cpp:181-187), and equals v_k bit
  for bit; the residual column likewise (ACA.cpp:188-195), and u_k = col / pivot;
* J_k is the first maximum of |residual row k| over the columns not yet used
  (maxAbsVector, ACA.cpp:10-23), and I_{k+1} the first maximum of the residual
  column k over the rows not yet used (ACA.cpp:213), not of the raw column;
* the rank is the first k at which ||u_k|| ||v_k|| > tol S_k no longer holds
  the loop open (ACA.cpp:178), with S_0 = 0 (the first pair never enters S) and
  S_k^2 = S_{k-1}^2 + (||u_k|| ||v_k||)^2 + sum_{l<=k} 2 |u_l.u_k| |v_l.v_k|,
  the sum including l = k (ACA.cpp:206-212); or min(n, m); or the step after
  the last pair produced a pair of norm below 1e-16 (ACA.cpp:196-200).

Each alternative reading (pivot rows from the raw column, S without the l = k
term, S that counts the first pair) is evaluated as well and must disagree with
the oracle on at least one block, so the checks can tell them apart.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import lattice, uniform


def dense(P, rs, n, cs, m):
    """1/r entries with the oracle's arithmetic: ((dx*dx + dy*dy) + dz*dz), IEEE sqrt, 1/R."""
    a = P[rs:rs + n][:, None, :]
    b = P[cs:cs + m][None, :, :]
    dx, dy, dz = a[..., 0] - b[..., 0], a[..., 1] - b[..., 1], a[..., 2] - b[..., 2]
    return 1.0 / np.sqrt((dx * dx + dy * dy) + dz * dz)


def first_max(v, used):
    """maxAbsVector: the first index of the largest |v| among indices not in `used`."""
    allowed = np.ones(len(v), bool)
    allowed[np.asarray(used, dtype=np.int64)] = False
    idx = np.flatnonzero(allowed)
    return int(idx[np.argmax(np.abs(v[idx]))])  # np.argmax returns the first maximum


def residual_row(A, U, V, i, k):
    row = A[i].copy()
    for l in range(k):
        row -= U[l, i] * V[l]
    return row


def residual_col(A, U, V, j, k):
    col = A[:, j].copy()
    for l in range(k):
        col -= V[l, j] * U[l]
    return col


def sample_blocks(o, level, nbox, seed, stride=3):
    s, c, _ = o.boxes(level)
    off, idx = o.il(level)
    rng = np.random.default_rng(seed)
    out = []
    for k in rng.choice(8 ** level, nbox, replace=False):
        for b in range(off[k], off[k + 1], stride):
            ki = idx[b]
            if c[k] and c[ki]:
                out.append((int(s[k]), int(c[k]), int(s[ki]), int(c[ki])))
    return out


CASES = {
    # name: (points, leaf1d, tol_pow, level, boxes sampled)
    "lattice16-tol10": (lambda: lattice(16), 4, 10, 2, 24),
    "lattice12-tol6": (lambda: lattice(12), 3, 6, 2, 24),
    "uniform4k-tol8": (lambda: uniform(4096, 31), 4, 8, 2, 24),
    "uniform8k-tol5": (lambda: uniform(8000, 32), 4, 5, 2, 24),
}


@pytest.fixture(scope="module", params=list(CASES))
def runs(request):
    mk, leaf, tol, level, nbox = CASES[request.param]
    x = mk()
    o = O.Oracle(x, leaf, tol_pow=tol, keep_uv=True)
    P = x[o.perm()]
    out = []
    for rs, n, cs, m in sample_blocks(o, level, nbox, seed=tol):
        r, I, J, U, V = o.aca_positions(rs, n, cs, m)
        out.append((dense(P, rs, n, cs, m), r, I, J, U, V))
    return request.param, 10.0 ** (-tol), out


def test_pivots_and_vectors_follow_the_residuals(runs):
    name, tol, out = runs
    raw_differs = 0
    for A, r, I, J, U, V in out:
        n, m = A.shape
        assert r >= 1
        # first row: the first row whose largest entry exceeds 1e-16 (ACA.cpp:152-161)
        assert I[0] == next(l for l in range(n) if np.abs(A[l]).max() > 1e-16)
        for k in range(r):
            row = residual_row(A, U, V, I[k], k)
            assert np.array_equal(row, V[k]), (name, k)
            assert J[k] == first_max(row, J[:k]), (name, k)
            col = residual_col(A, U, V, J[k], k)
            assert np.array_equal(col / row[J[k]], U[k]), (name, k)
            if k + 1 < r:
                assert I[k + 1] == first_max(col, I[:k + 1]), (name, k)
                raw_differs += I[k + 1] != first_max(A[:, J[k]], I[:k + 1])
    # reading the pivot row off the raw column would pick other rows on these blocks
    assert raw_differs > 0, name


def predicted_rank(A, U, V, I, J, r, tol, variant):
    """Rank given by the stop rule `variant` from the returned pairs, or None
    when one of its decisions lies within 1e-9 (relative) of the threshold
    (numpy's norms sum in another order than the oracle's)."""
    n, m = A.shape
    mn = min(n, m)
    nu = [np.linalg.norm(U[k]) for k in range(r)]
    nv = [np.linalg.norm(V[k]) for k in range(r)]
    S = nu[0] * nv[0] if variant == "first-pair-in-S" else 0.0
    for k in range(r):
        if k >= 1:
            s2 = S * S + (nu[k] * nv[k]) ** 2
            for l in range(k if variant == "no-self-term" else k + 1):
                s2 += 2 * abs(U[l] @ U[k]) * abs(V[l] @ V[k])
            S = np.sqrt(s2)
        lhs, rhs = nu[k] * nv[k], tol * S
        if rhs > 0 and abs(lhs - rhs) <= 1e-9 * rhs:
            return None
        if not lhs > rhs or k + 1 >= mn:
            return k + 1
    # the loop was still open after the last returned pair: the rank stays r
    # only if the next candidate pair has a norm below 1e-16 (ACA.cpp:196-200)
    col = residual_col(A, U, V, J[r - 1], r - 1)
    i = first_max(col, I[:r])
    row = residual_row(A, U, V, i, r)
    j = first_max(row, J[:r])
    u = residual_col(A, U, V, j, r) / row[j]
    if np.linalg.norm(u) < 1e-16 or np.linalg.norm(row) < 1e-16:
        return r
    return r + 1  # at least one more pair


def test_stop_rule_includes_the_new_pair_in_S(runs):
    name, tol, out = runs
    agree = {"reference": 0, "no-self-term": 0, "first-pair-in-S": 0}
    differ = dict.fromkeys(agree, 0)
    for A, r, I, J, U, V in out:
        for variant in agree:
            p = predicted_rank(A, U, V, I, J, r, tol, variant)
            if p is None:
                continue
            if p == r:
                agree[variant] += 1
            else:
                differ[variant] += 1
    assert differ["reference"] == 0 and agree["reference"] >= 0.95 * len(out), (name, agree,
                                                                                  differ)
    print(f"[{name}] blocks {len(out)} agree {agree} differ {differ}")


def test_alternative_stop_rules_are_distinguished():
    """Over all cases, each misreading of the normS update changes at least one
    rank, so the stop-rule test above would catch it."""
    differ = {"no-self-term": 0, "first-pair-in-S": 0}
    for name, (mk, leaf, tol, level, nbox) in CASES.items():
        x = mk()
        o = O.Oracle(x, leaf, tol_pow=tol, keep_uv=True)
        P = x[o.perm()]
        for rs, n, cs, m in sample_blocks(o, level, nbox, seed=tol):
            r, I, J, U, V = o.aca_positions(rs, n, cs, m)
            A = dense(P, rs, n, cs, m)
            for variant in differ:
                p = predicted_rank(A, U, V, I, J, r, 10.0 ** (-tol), variant)
                differ[variant] += p is not None and p != r
    assert all(v > 0 for v in differ.values()), differ


def test_find_memory_and_max_rank_from_the_structure():
    """findMemory (src/HODLR3DTree.cpp:1757-1778) counted from the exported
    ranks, box sizes, lists and neighbours: r n_k + r^2 + m r per block, plus
    n_k^2 + n_k n_nn per leaf; getMaxRank (:1807-1818) is the largest rank."""
    x = uniform(6000, 33)
    o = O.Oracle(x, 4, tol_pow=8)
    o.assemble()
    total, mr = 0, 0
    for j in range(1, o.nLevels + 1):
        _, c, _ = o.boxes(j)
        off, idx = o.il(j)
        r, _ = o.ranks(j)
        tgt = np.repeat(np.arange(8 ** j), np.diff(off))
        n, m = c[tgt].astype(np.int64), c[idx].astype(np.int64)
        rr = r.astype(np.int64)
        total += int((rr * n + rr * rr + m * rr).sum())
        mr = max(mr, int(r.max(initial=0)))
    _, c, _ = o.boxes(o.nLevels)
    nb = o.neighbors(o.nLevels)
    c64 = c.astype(np.int64)
    total += int((c64 * c64).sum())
    total += int(sum(c64[k] * c64[nn] for k in range(len(c)) for nn in nb[k] if nn >= 0))
    assert o.find_memory() == float(total)
    assert o.max_rank() == mr
