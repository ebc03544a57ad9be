"""Oracle tree, bucketing and interaction lists (CPU).

Pins: the reference's interaction-list push order extracted from its source
text (tests/golden/il_order.json), the nLevels values of the reference formula,
closed-form geometry (Chebyshev/Manhattan classification), per-level block
counts, and bucketing invariants.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import lattice, uniform

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "il_order.json")


def slot_offset(s):
    m = s if s < 13 else s + 1
    return (m % 3 - 1, (m % 9) // 3 - 1, m // 9 - 1)


def child_bits(c):
    return (int(c in (1, 2, 5, 6)), int(c in (2, 3, 6, 7)), int(c >= 4))


def box_coords(k, j):
    X = Y = Z = 0
    for t in range(j):
        c = (k >> (3 * (j - 1 - t))) & 7
        b = child_bits(c)
        X, Y, Z = 2 * X + b[0], 2 * Y + b[1], 2 * Z + b[2]
    return X, Y, Z


def test_nlevels_table():
    # src/HODLR3D.cpp:6-7 evaluated with glibc; values listed for the configs
    cases = {(4096, 4): 2, (262144, 4): 4, (2 ** 20, 4): 5, (2 ** 21, 4): 5, (8192, 4): 3,
             (32768, 4): 3, (32769, 4): 4, (200000, 4): 4, (8000, 6): 2, (729000, 6): 4,
             (1000000, 6): 5}
    for (N, l), v in cases.items():
        assert O.nlevels(N, l) == v, (N, l)


def test_il_counts_per_level():
    o = O.Oracle(uniform(300, 0), 4, nlevels=5)
    counts = [len(o.il(j)[1]) for j in range(1, 6)]
    assert counts == [8, 2800, 42272, 435520, 3909248]
    nbr = [(o.neighbors(j) >= 0).sum() for j in range(2, 6)]
    assert nbr == [720, 7392, 66240, 559488]
    for j in range(2, 6):
        off, _ = o.il(j)
        d = np.diff(off)
        assert d.min() == 37 and d.max() == (49 if j == 2 else 133)


def expected_il(table, c, parent_slots_present, k):
    """IL of child c of parent p = k//8 from the fixture, given which parent slots exist
    (each parent-slot block is guarded by parent.neighborNumbers[P] != -1)."""
    out = []
    for ev in table[str(c)]:
        if ev[0] == "sib_il":
            out.append(("sib", ev[1]))
        elif ev[0] == "par" and ev[1] in parent_slots_present:
            out.extend(("par", ev[1], op[1]) for op in ev[2] if op[0] == "il")
    return out


@pytest.mark.parametrize("level", [2, 3, 4])
def test_il_order_matches_reference_text(level):
    table = json.load(open(GOLDEN))["children"]
    o = O.Oracle(uniform(100, 1), 4, nlevels=level)
    off, idx = o.il(level)
    pn = o.neighbors(level - 1)
    rng = np.random.default_rng(level)
    boxes = range(8 ** level) if level <= 3 else rng.choice(8 ** level, 800, replace=False)
    for k in boxes:
        p = k // 8
        present = {s: pn[p, s] for s in range(26) if pn[p, s] >= 0}
        exp = []
        for e in expected_il(table, k % 8, present, k):
            if e[0] == "sib":
                exp.append(8 * p + e[1])
            else:
                exp.append(8 * present[e[1]] + e[2])
        assert list(idx[off[k]:off[k + 1]]) == exp, k


def test_neighbor_slots_fixture_and_geometry():
    table = json.load(open(GOLDEN))["children"]
    for c in range(8):
        P = child_bits(c)
        for ev in table[str(c)]:
            if ev[0] == "sib_nbr":
                d = tuple(a - b for a, b in zip(child_bits(ev[2]), P))
                assert slot_offset(ev[1]) == d
    o = O.Oracle(uniform(100, 2), 4, nlevels=3)
    n = o.neighbors(3)
    for k in range(8 ** 3):
        X = box_coords(k, 3)
        for s in range(26):
            d = slot_offset(s)
            inside = all(0 <= X[i] + d[i] < 8 for i in range(3))
            if sum(map(abs, d)) <= 2 and inside:
                assert box_coords(n[k, s], 3) == tuple(X[i] + d[i] for i in range(3))
            else:
                assert n[k, s] == -1


def test_pairs_partitioned_exactly_once():
    """Every ordered pair of leaves is covered by exactly one of: near (self, face,
    edge) or one interaction-list block at one level (weak admissibility)."""
    L = 3
    o = O.Oracle(uniform(100, 3), 4, nlevels=L)
    nleaf = 8 ** L
    cover = np.zeros((nleaf, nleaf), dtype=np.int32)
    n = o.neighbors(L)
    for k in range(nleaf):
        cover[k, k] += 1
        for s in range(26):
            if n[k, s] >= 0:
                cover[k, n[k, s]] += 1
    for j in range(1, L + 1):
        off, idx = o.il(j)
        span = 8 ** (L - j)
        for k in range(8 ** j):
            for ki in idx[off[k]:off[k + 1]]:
                cover[k * span:(k + 1) * span, ki * span:(ki + 1) * span] += 1
    assert cover.min() == 1 and cover.max() == 1


def test_il_is_admissible_and_complete():
    o = O.Oracle(uniform(100, 4), 4, nlevels=4)
    for j in (2, 3, 4):
        off, idx = o.il(j)
        C = np.array([box_coords(q, j) for q in range(8 ** j)])
        P = C >> 1
        for k in np.random.default_rng(j).choice(8 ** j, min(300, 8 ** j), replace=False):
            got = set(idx[off[k]:off[k + 1]])
            # children of the parent's neighbours and siblings that are not self, face or edge
            dp = np.abs(P - P[k])
            d = np.abs(C - C[k])
            parent_ok = (dp.max(1) <= 1) & (dp.sum(1) < 3)
            near = (d.max(1) == 0) | ((d.max(1) == 1) & (d.sum(1) <= 2))
            exp = set(np.nonzero(parent_ok & ~near)[0])
            assert got == exp


@pytest.mark.parametrize("pts", [lattice(12), uniform(5000, 5), uniform(3000, 6, -0.3, 0.9)])
def test_bucketing_invariants(pts):
    o = O.Oracle(pts, 3)
    L = o.nLevels
    perm = o.perm()
    assert sorted(perm) == list(range(len(pts)))
    s, c, ctr = o.boxes(L)
    h = 1.0 / 2 ** L
    for k in range(8 ** L):
        seg = perm[s[k]:s[k] + c[k]]
        assert list(seg) == sorted(seg)  # ascending original index within a leaf
        p = pts[seg]
        lo, hi = ctr[k] - h, ctr[k] + h
        # `<=` rule: a point on a split plane belongs to the lower child
        interior = ~((np.abs(np.abs(pts[seg]) - 1) < 0) )
        assert np.all(p <= hi + 1e-15) or not interior.any()
        inner = (p > lo) | (np.isclose(lo, -1.0))
        assert np.all(inner | (p < lo))  # points below the box only if outside the root
    for j in range(1, L):
        sj, cj, _ = o.boxes(j)
        sc, cc, _ = o.boxes(j + 1)
        assert np.array_equal(sj, sc[::8])
        assert np.array_equal(cj, cc.reshape(-1, 8).sum(1))


def test_split_plane_ties_go_to_lower_child():
    # a point exactly on the root centre planes lands in child 0 at level 1
    pts = np.array([[0.0, 0.0, 0.0], [0.5, -0.5, 0.5], [1e-300, 0, 0]] + [[0.9, 0.9, 0.9]] * 0)
    pts = np.vstack([pts, uniform(61, 9)])
    o = O.Oracle(pts, 1, nlevels=1)
    s, c, _ = o.boxes(1)
    perm = list(o.perm())
    assert perm.index(0) < c[0]  # (0,0,0) -> child 0
    leaf_of_2 = np.searchsorted(np.cumsum(c), perm.index(2), side="right")
    assert leaf_of_2 == 1  # x > 0 only -> child 1 (+,-,-)
