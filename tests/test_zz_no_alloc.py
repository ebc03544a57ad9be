"""h3d_matvec_noalloc, the allocation-free product (GPU): through the C ABI,
bit for bit equal to h3d_matvec in every resident mode, on two handles in a
row (the second with fewer points, sharing the per-GPU scratch array), then
the first again; oversized products and non-resident precisions refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def noalloc(h, b, prec):
    import torch
    from hodlr3d_amd import _native as nat
    from hodlr3d_amd.hodlr3d import _prec
    y = torch.full_like(b, float("nan"))
    torch.cuda.synchronize()
    rc = nat.lib().h3d_matvec_noalloc(h._h, ctypes.c_void_p(b.data_ptr()),
                                      ctypes.c_void_p(y.data_ptr()), _prec(prec))
    nat.check(rc)
    nat.check(nat.lib().h3d_synchronize(h._h))
    return y


def handle(N, seed, tol=8):
    import hodlr3d_amd as H3
    x = np.random.default_rng(seed).uniform(-1, 1, (N, 3))
    h = H3.HODLR3D(4, 1.0, tol, 7, x)
    h.assemble()
    return h


def test_noalloc_product_equals_the_allocating_one():
    import torch
    a = handle(30000, 91)
    a.prepare_precision("fp32")
    a.compress()
    gen = torch.Generator(device="cuda").manual_seed(92)
    ba = torch.randn(a.N, dtype=torch.float64, device="cuda", generator=gen)
    modes = ("fp64", "fp32", "fp32/acc64", "mixed")
    ref = {m: a.matvec(ba, prec=m) for m in modes}
    for m in modes:
        print(f"[noalloc] N={a.N} {m}", flush=True)
        assert torch.equal(noalloc(a, ba, m), ref[m]), m
    # a second handle with fewer points: same shared array, smaller carve
    b = handle(9000, 93)
    bb = torch.randn(b.N, dtype=torch.float64, device="cuda", generator=gen)
    print(f"[noalloc] N={b.N} fp64", flush=True)
    assert torch.equal(noalloc(b, bb, "fp64"), b.matvec(bb))
    # and the first again, after the smaller call left its contents
    for m in modes:
        assert torch.equal(noalloc(a, ba, m), ref[m]), m
    from hodlr3d_amd import _native as nat
    with pytest.raises(nat.H3DError) as ei:
        noalloc(b, bb, "bf16")  # no bf16 copy resident: nothing is built here
    assert ei.value.code == nat.H3D_ERR_STATE
    a.close()
    b.close()


def test_noalloc_refuses_products_beyond_its_scratch():
    import torch
    from hodlr3d_amd import _native as nat
    c = handle(200000, 94)
    bc = torch.randn(c.N, dtype=torch.float64, device="cuda")
    with pytest.raises(nat.H3DError) as ei:
        noalloc(c, bc, "fp64")
    assert ei.value.code == nat.H3D_ERR_INVALID and "scratch" in str(ei.value)
    c.close()
