"""Batched products of several independent instances (each with its own root
cube) captured into one HIP graph: every replay must equal the eager products
bit for bit, with eager work and synchronisation between the replays (GPU)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def instances():
    import hodlr3d_amd as H3
    rng = np.random.default_rng(31)
    hs = []
    for n in (3000, 6000, 9000, 12000):
        half = rng.uniform(0.25, 1.0)           # box side in [0.5, 2]
        c = rng.uniform(-3.0, 3.0, 3)
        pts = c + rng.uniform(-half, half, (n, 3))
        h = H3.HODLR3D(4, half, 8, 7, pts, center=c)
        h.assemble()
        print(f"[instance] N={n} centre={np.round(c, 3)} half-width={half:.3f} "
              f"levels={h.nLevels}", flush=True)
        hs.append(h)
    yield hs
    for h in hs:
        h.close()


@pytest.mark.parametrize("prec,transpose", [("fp64", False), ("fp64", True), ("mixed", False)])
def test_graph_replays_equal_eager_products(instances, prec, transpose):
    import torch
    import hodlr3d_amd as H3
    hs = instances
    if prec == "mixed":
        for h in hs:
            h.compress(1e-7)
    gen = torch.Generator(device="cuda").manual_seed(5)
    b = [torch.randn(h.N, dtype=torch.float64, device="cuda", generator=gen) for h in hs]
    y = [torch.empty(h.N, dtype=torch.float64, device="cuda") for h in hs]
    g = H3.MatvecGraph(hs, b, y, prec=prec, transpose=transpose)
    for rep in range(4):
        if rep:
            for bi in b:
                bi.copy_(torch.randn(bi.shape, dtype=torch.float64, device="cuda", generator=gen))
        ref = [h.matvec(bi, prec=prec, transpose=transpose) for h, bi in zip(hs, b)]
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        for i, (yi, ri) in enumerate(zip(y, ref)):
            assert torch.equal(yi, ri), (prec, transpose, rep, i)
    g.close()


def test_repeated_handle_keeps_its_products_in_order(instances):
    """The same handle twice in one graph: both products share its scratch,
    so they must run in order (one branch) and equal the eager products."""
    import torch
    import hodlr3d_amd as H3
    hs = instances
    seq = [hs[0], hs[1], hs[0], hs[2], hs[3], hs[1]]
    gen = torch.Generator(device="cuda").manual_seed(9)
    b = [torch.randn(h.N, dtype=torch.float64, device="cuda", generator=gen) for h in seq]
    y = [torch.empty(h.N, dtype=torch.float64, device="cuda") for h in seq]
    g = H3.MatvecGraph(seq, b, y, prec="fp64")
    for rep in range(3):
        ref = [h.matvec(bi) for h, bi in zip(seq, b)]
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        for i, (yi, ri) in enumerate(zip(y, ref)):
            assert torch.equal(yi, ri), (rep, i)
    g.close()


def test_graph_rejects_bad_input(instances):
    import torch
    import hodlr3d_amd as H3
    hs = instances
    b = [torch.zeros(hs[0].N + 1, dtype=torch.float64, device="cuda")]
    y = [torch.zeros(hs[0].N, dtype=torch.float64, device="cuda")]
    with pytest.raises(ValueError):
        H3.MatvecGraph(hs[:1], b, y)
