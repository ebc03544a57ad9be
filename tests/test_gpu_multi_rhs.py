"""Products with several right-hand sides (the grouped matrix-core path):
every column equals the single-RHS product of the same mode within that
mode's rounding, for group widths 16 and 32, ragged counts, host and device
buffers; a non-finite column stays in its own column (GPU).

fp64 and the fp64-accumulation modes sum in fp64 on both paths (different
order): 1e-13.  fp32/bf16 with fp32 sums: 1e-5 relative (fp32 rounding of
~1e3-term sums in another order)."""
import numpy as np
import pytest

from tests.conftest import uniform

pytestmark = pytest.mark.gpu

MODES = {"fp64": 1e-13, "fp32/acc64": 1e-13, "bf16/acc64": 1e-13, "fp32": 1e-5, "bf16": 1e-5}


@pytest.fixture(scope="module")
def handle():
    import hodlr3d_amd as H3
    x = uniform(24000, 51)
    h = H3.HODLR3D(4, 1.0, 8, 7, x)
    h.assemble()
    print(f"[multi-rhs] N={len(x)} levels={h.nLevels}", flush=True)
    yield x, h
    h.close()


def _single(h, Q, prec):
    return np.stack([h.matvec(Q[:, j], prec=prec) for j in range(Q.shape[1])], axis=1)


def _relcols(A, B):
    return max(np.linalg.norm(A[:, j] - B[:, j]) / np.linalg.norm(B[:, j]) for j in range(B.shape[1]))


@pytest.mark.parametrize("prec", list(MODES))
def test_grouped_columns_equal_single_products(handle, prec):
    x, h = handle
    tol = MODES[prec]
    Q = np.random.default_rng(1).standard_normal((len(x), 40))
    ref = _single(h, Q, prec)
    h.set_rhs_block(32)
    for n in (2, 5, 16, 17, 40):
        print(f"[{prec}] nrhs={n}", flush=True)
        Y = h.matvec(Q[:, :n], prec=prec)
        assert _relcols(Y, ref[:, :n]) <= tol, (prec, n)
    h.set_rhs_block(16)
    Y = h.matvec(Q[:, :40], prec=prec)
    assert _relcols(Y, ref) <= tol, (prec, "width 16")
    h.set_rhs_block(32)


def test_width_zero_is_the_single_path(handle):
    x, h = handle
    Q = np.random.default_rng(2).standard_normal((len(x), 3))
    h.set_rhs_block(0)
    Y = h.matvec(Q)
    h.set_rhs_block(32)
    assert np.array_equal(Y, _single(h, Q, "fp64"))


def test_device_buffers_and_direct_rows(handle):
    import torch
    x, h = handle
    N = len(x)
    gen = torch.Generator(device="cuda").manual_seed(3)
    Qt = torch.randn(N, 20, dtype=torch.float64, device="cuda", generator=gen)
    Yt = h.matvec(Qt)
    Yh = h.matvec(Qt.cpu().numpy())
    assert np.linalg.norm(Yt.cpu().numpy() - Yh) <= 1e-15 * np.linalg.norm(Yh)
    rows = np.arange(0, N, 97)
    for j in (0, 19):
        d = h.direct(Qt[:, j].cpu().numpy(), rows)
        assert np.linalg.norm(Yh[rows, j] - d) <= 1e-6 * np.linalg.norm(d), j


def test_nonfinite_column_stays_in_its_column(handle):
    x, h = handle
    Q = np.random.default_rng(4).standard_normal((len(x), 6))
    Q[123, 2] = np.nan
    Y = h.matvec(Q)
    assert np.isnan(Y[:, 2]).any()
    good = [0, 1, 3, 4, 5]
    assert np.isfinite(Y[:, good]).all()
    assert _relcols(Y[:, good], _single(h, Q[:, good], "fp64")) <= 1e-13


@pytest.mark.parametrize("prec", list(MODES))
def test_grouped_transposes_equal_single_transposes(handle, prec):
    """A^T with several right-hand sides: phase 1 takes U_k^T x_k per target
    box, phase 2 sums V_b t_b over the blocks of each source box (columns from
    a table); every column equals the single transposed product within the
    mode's rounding, for widths 16 and 32 and ragged counts."""
    x, h = handle
    tol = MODES[prec]
    Q = np.random.default_rng(9).standard_normal((len(x), 40))
    ref = np.stack([h.matvec(Q[:, j], prec=prec, transpose=True) for j in range(40)], axis=1)
    for width, counts in ((32, (2, 17, 40)), (16, (16, 40))):
        h.set_rhs_block(width)
        for n in counts:
            print(f"[{prec} transposed] width {width} nrhs={n}", flush=True)
            Y = h.matvec(Q[:, :n], prec=prec, transpose=True)
            assert _relcols(Y, ref[:, :n]) <= tol, (prec, width, n)
    h.set_rhs_block(32)


def test_grouped_transpose_is_the_adjoint(handle):
    """<A^T X, Z> = <X, A Z> column by column, both sides grouped (fp64)."""
    x, h = handle
    rng = np.random.default_rng(10)
    X = rng.standard_normal((len(x), 20))
    Z = rng.standard_normal((len(x), 20))
    AtX = h.matvec(X, transpose=True)
    AZ = h.matvec(Z)
    lhs = np.einsum("ij,ij->j", AtX, Z)
    rhs = np.einsum("ij,ij->j", X, AZ)
    assert np.max(np.abs(lhs - rhs) / np.abs(rhs)) <= 1e-12


def test_grouped_mixed_products(handle):
    """The recompressed copy in groups: each class through the grouped phases
    (fp32 factors with fp64 sums, fp16 factors with fp32 sums, scaled t), the
    near field with fp64 coordinates and sums.  Columns agree with the single
    mixed products (whose near field forms fp32 entries: ~6e-8 of y apart) and
    stay within the copy's budget of the direct sum; repeated calls give the
    same bits, widths 16 and 32 the same columns."""
    x, h = handle
    N = len(x)
    Q = np.random.default_rng(14).standard_normal((N, 20))
    tol = 1e-6
    h.compress(tol)
    ref = np.stack([h.matvec(Q[:, j], prec="mixed") for j in range(20)], axis=1)
    h.set_rhs_block(32)
    Y = h.matvec(Q, prec="mixed")
    assert _relcols(Y, ref) <= 2e-7
    assert np.array_equal(h.matvec(Q, prec="mixed"), Y)
    h.set_rhs_block(16)
    assert np.array_equal(h.matvec(Q, prec="mixed"), Y)
    h.set_rhs_block(32)
    rows = np.arange(0, N, 61)
    for j in (0, 19):
        d = h.direct(Q[:, j], rows)
        assert np.linalg.norm(Y[rows, j] - d) <= 3 * tol * np.linalg.norm(d), j
    h.release_precision("mixed")


def test_mixed_transposes_stay_per_column(handle):
    x, h = handle
    Q = np.random.default_rng(5).standard_normal((len(x), 3))
    h.compress(1e-6)
    YM = h.matvec(Q, prec="mixed", transpose=True)
    for j in range(3):
        assert np.array_equal(YM[:, j], h.matvec(Q[:, j], prec="mixed", transpose=True))
    h.release_precision("mixed")


def test_rejects_bad_width(handle):
    from hodlr3d_amd import _native as nat
    x, h = handle
    with pytest.raises(nat.H3DError):
        h.set_rhs_block(8)


def test_clustered_leaves_with_many_rows():
    """A leaf with ~700 points (several source windows, and row-split items in
    the single-RHS plan) next to ordinary leaves."""
    import hodlr3d_amd as H3
    rng = np.random.default_rng(7)
    base = rng.uniform(-1, 1, (6000, 3))
    ball = np.array([0.31, -0.42, 0.55]) + rng.uniform(-0.04, 0.04, (700, 3))
    x = np.vstack([base, ball])
    h = H3.HODLR3D(3, 1.0, 9, 7, x)
    h.assemble()
    Q = np.random.default_rng(8).standard_normal((len(x), 20))
    for prec, tol in (("fp64", 1e-13), ("fp32", 1e-5)):
        ref = _single(h, Q, prec)
        Y = h.matvec(Q, prec=prec)
        assert _relcols(Y, ref) <= tol, prec
    h.close()


def test_autotune_rhs_keeps_the_product(handle):
    """The group width and the phase-1 item size change only the schedule:
    each column's sums keep their order (matrix-core products over the same
    4-row steps, the same per-column near-field loop), so products are
    bit-identical whatever the autotuner picks."""
    x, h = handle
    Q = np.random.default_rng(12).standard_normal((len(x), 24))
    refs = {}
    for w in (16, 32):
        h.set_rhs_block(w)
        refs[w] = {p: h.matvec(Q, prec=p) for p in ("fp32", "fp64", "bf16")}
    for p in ("fp32", "fp64", "bf16"):
        assert np.array_equal(refs[16][p], refs[32][p]), p
    res = h.autotune_rhs("fp32", nrhs=24, reps=1)
    assert res["width"] in (16, 32) and res["columns"] in (1024, 2048)
    assert len(res["ms"]) == 4 and all(v > 0 for v in res["ms"].values())
    assert np.array_equal(h.matvec(Q, prec="fp32"), refs[32]["fp32"])
    for w in (16, 32):
        h.set_rhs_block(w)
        for p in ("fp32", "fp64"):
            assert np.array_equal(h.matvec(Q, prec=p), refs[w][p]), (w, p)
    h.set_rhs_block(32)
    ref = h.matvec(Q, prec="fp32")
    res = h.autotune_rhs("fp32", nrhs=24, reps=1)
    assert res["width"] in (16, 32) and res["columns"] in (1024, 2048)
    assert len(res["ms"]) == 4 and all(v > 0 for v in res["ms"].values())
    assert np.array_equal(h.matvec(Q, prec="fp32"), ref)
    for w in (16, 32):
        h.set_rhs_block(w)
        assert np.array_equal(h.matvec(Q, prec="fp32"), ref), w
    h.set_rhs_block(32)


@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp64", "fp32/acc64"])
def test_grouped_products_repeat_bit_for_bit(handle, prec):
    """A grouped product is deterministic: the same call gives the same bits.
    The 16-wide fp32-sum phase-1 kernel once stored matrix-core accumulators
    before the last product had completed (blocks with at most 16 columns
    left in a pass), which showed only as run-to-run differences of ~1e-8."""
    x, h = handle
    Q = np.random.default_rng(13).standard_normal((len(x), 24))
    for w in (16, 32):
        h.set_rhs_block(w)
        for tr in (False, True):
            for cnt in (3, 16, 24):
                first = h.matvec(Q[:, :cnt], prec=prec, transpose=tr)
                for _ in range(3):
                    assert np.array_equal(h.matvec(Q[:, :cnt], prec=prec, transpose=tr), first), \
                        (prec, w, tr, cnt)
    h.set_rhs_block(32)
