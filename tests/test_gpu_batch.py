"""The instance batch (GPU): the work tables of many handles merged into one
launch per step (h3d_batch_create).  Every instance's slice of a batch product
must equal that handle's own product bit for bit at the same load widths, for
each ACA copy and with per-instance operators; graph replays must equal the
eager batch product; products must sit within the fp32 error of the direct sum.
The caller side batched here is HODLR3D::MatVecProduct (src/HODLR3D.cpp:21-27).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_instances(Ns, seed, store=("fp64",), q=7):
    import hodlr3d_amd as H3
    rng = np.random.default_rng(seed)
    hs = []
    for n in Ns:
        half = rng.uniform(0.25, 1.0)           # box side in [0.5, 2]
        c = rng.uniform(-3.0, 3.0, 3)
        pts = c + rng.uniform(-half, half, (n, 3))
        h = H3.HODLR3D(4, half, 8, q, pts, center=c, store=store)
        h.assemble()
        print(f"[batch instance] N={n} levels={h.nLevels} q={q}", flush=True)
        hs.append(h)
    return hs


@pytest.fixture(scope="module")
def instances():
    hs = make_instances((3000, 9000, 600, 20000, 5000), seed=51)
    yield hs
    for h in hs:
        h.close()


def packed_rhs(bt, seed):
    import torch
    gen = torch.Generator(device="cuda").manual_seed(seed)
    return torch.randn(bt.N, dtype=torch.float64, device="cuda", generator=gen)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp32/acc64", "bf16", "bf16/acc64"])
def test_slices_equal_each_handles_own_product(instances, prec):
    import torch
    import hodlr3d_amd as H3
    hs = instances
    for h in hs:
        h.set_tuning(prec, 16, 16)
    bt = H3.Batch(hs, prec=prec)
    assert bt.get_tuning() == (16, 16)
    assert list(np.diff(bt.offsets)) == [h.N for h in hs]
    b = packed_rhs(bt, 1)
    y = bt.matvec(b)
    torch.cuda.synchronize()
    for i, (h, bi, yi) in enumerate(zip(hs, bt.unpack(b), bt.unpack(y))):
        ref = h.matvec(bi.contiguous(), prec=prec)
        assert torch.equal(yi, ref), (prec, i, float((yi - ref).abs().max()))
    bt.close()


def test_per_instance_operators(instances):
    """alpha, a scalar diagonal and a per-point diagonal differ per instance."""
    import torch
    import hodlr3d_amd as H3
    hs = instances
    try:
        for k, h in enumerate(hs):
            h.set_operator(alpha=1.0 + 0.25 * k, diag=0.5 * k)
        c = np.random.default_rng(3).uniform(1.0, 1.1, hs[1].N)
        hs[1].set_diag(c)
        for h in hs:
            h.set_tuning("fp32", 16, 16)
        bt = H3.Batch(hs, prec="fp32")
        b = packed_rhs(bt, 2)
        y = bt.matvec(b)
        for i, (h, bi, yi) in enumerate(zip(hs, bt.unpack(b), bt.unpack(y))):
            assert torch.equal(yi, h.matvec(bi.contiguous(), prec="fp32")), i
        bt.close()
    finally:
        hs[1].set_diag(None)
        for h in hs:
            h.set_operator(alpha=1.0, diag=0.0)


def test_graph_replays_equal_eager_batch_products(instances):
    import torch
    import hodlr3d_amd as H3
    hs = instances
    bt = H3.Batch(hs, prec="fp32")
    b = packed_rhs(bt, 4)
    y = torch.empty_like(b)
    launch = bt.graph(b, y)
    gen = torch.Generator(device="cuda").manual_seed(6)
    for rep in range(4):
        if rep:
            b.copy_(torch.randn(b.shape, dtype=torch.float64, device="cuda", generator=gen))
        ref = bt.matvec(b)
        torch.cuda.synchronize()
        launch()
        torch.cuda.synchronize()
        assert torch.equal(y, ref), rep
        _ = [h.matvec(bi.contiguous(), prec="fp32") for h, bi in zip(hs, bt.unpack(b))]
    bt.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_persistent_launch_equals_one_wave_per_item(instances, prec):
    """The work-queue launch takes the same items in another order: every
    item's sums are its own, so the product is the same bit for bit, on
    repeated launches (the queue counters are cleared before each) and in a
    graph replay."""
    import torch
    import hodlr3d_amd as H3
    hs = instances
    bt = H3.Batch(hs, prec=prec)
    b = packed_rhs(bt, 9)
    ref = bt.matvec(b)
    for p1, p2 in ((H3.PERSISTENT, 0), (0, H3.PERSISTENT), (H3.PERSISTENT, H3.PERSISTENT)):
        bt.set_tuning(16 | p1, 16 | p2)
        assert bt.get_tuning() == (16 | p1, 16 | p2)
        for rep in range(3):
            y = bt.matvec(b)
            torch.cuda.synchronize()
            assert torch.equal(y, ref), (prec, p1, p2, rep)
    yg = torch.empty_like(b)
    launch = bt.graph(b, yg)
    for rep in range(3):
        launch()
        torch.cuda.synchronize()
        assert torch.equal(yg, ref), (prec, rep)
    bt.close()


def test_autotune_keeps_the_product(instances):
    import torch
    import hodlr3d_amd as H3
    hs = instances
    bt = H3.Batch(hs, prec="fp32")
    ms = bt.autotune(reps=2)
    assert set(ms["phase1"]) == {16, 8, 4} and set(ms["phase2"]) == {16, 8, 4}
    assert ms["phase1_persistent"] > 0 and ms["phase2_persistent"] > 0
    w1, w2 = bt.get_tuning()
    for h in hs:
        h.set_tuning("fp32", w1 & ~H3.PERSISTENT, w2 & ~H3.PERSISTENT)
    b = packed_rhs(bt, 7)
    y = bt.matvec(b)
    for h, bi, yi in zip(hs, bt.unpack(b), bt.unpack(y)):
        assert torch.equal(yi, h.matvec(bi.contiguous(), prec="fp32"))
    for h in hs:
        h.set_tuning("fp32", 16, 16)
    bt.close()


def test_against_the_direct_sum():
    import torch
    import hodlr3d_amd as H3
    hs = make_instances((8192, 30000, 70000), seed=52, store=("fp32",))
    try:
        bt = H3.Batch(hs, prec="fp32")
        b = packed_rhs(bt, 8)
        y = bt.matvec(b)
        for h, bi, yi in zip(hs, bt.unpack(b), bt.unpack(y)):
            rows = np.sort(np.random.default_rng(h.N).choice(h.N, 512, replace=False))
            d = h.direct(bi.contiguous(), rows).cpu().numpy()
            e = np.linalg.norm(yi.cpu().numpy()[rows] - d) / np.linalg.norm(d)
            print(f"[batch] N={h.N} fp32 vs direct {e:.2e}", flush=True)
            assert e <= 5e-6, (h.N, e)
        bt.close()
    finally:
        for h in hs:
            h.close()
        torch.cuda.empty_cache()


def test_bad_batches_are_refused(instances):
    import hodlr3d_amd as H3
    hs = instances
    with pytest.raises(H3.H3DError):
        H3.Batch([hs[0], hs[1], hs[0]], prec="fp32")   # a handle twice
    with pytest.raises(H3.H3DError):
        H3.Batch(hs, prec="mixed")                     # not an ACA copy
    other = make_instances((2000,), seed=53, q=3)
    try:
        with pytest.raises(H3.H3DError):
            H3.Batch([hs[0], other[0]], prec="fp64")   # different kernels
    finally:
        other[0].close()
    bt = H3.Batch(hs[:2], prec="fp64")
    with pytest.raises(H3.H3DError):
        bt.set_tuning(2, 16)                           # narrower than an fp64 element
    bt.close()
