"""Failure handling through the C ABI (GPU): errors come back as status codes."""
import numpy as np
import pytest

from tests.conftest import uniform

pytestmark = pytest.mark.gpu


def test_duplicate_points_are_reported():
    import hodlr3d_amd as H3
    from hodlr3d_amd import _native as nat
    x = uniform(1000, 0)
    x[10] = x[500]
    with pytest.raises(H3.H3DError) as e:
        H3.HODLR3D(2, 1.0, 7, 7, x)
    assert e.value.code == nat.H3D_ERR_NONFINITE


def test_call_order_and_arguments():
    import hodlr3d_amd as H3
    from hodlr3d_amd import _native as nat
    h = H3.HODLR3D(2, 1.0, 7, 7, uniform(1000, 1))
    with pytest.raises(H3.H3DError) as e:
        h.matvec(np.ones(1000))
    assert e.value.code == nat.H3D_ERR_STATE
    h.assemble()
    with pytest.raises(H3.H3DError):
        h.assemble()
    with pytest.raises(ValueError):
        h.matvec(np.ones(999))
    with pytest.raises(H3.H3DError):
        h.matvec(np.ones(1000), prec=7)
    assert np.all(np.isfinite(h.matvec(np.ones(1000))))
    h.close()


def test_nonfinite_rhs_propagates_without_fault():
    import hodlr3d_amd as H3
    h = H3.HODLR3D(2, 1.0, 7, 7, uniform(1000, 2))
    h.assemble()
    q = np.ones(1000)
    q[3] = np.nan
    y = h.matvec(q)
    assert np.isnan(y).any()
    h.close()


def test_torch_and_numpy_products_interleave_in_order():
    """Device-tensor products run on torch's stream and host-array products on
    the handle's own; both share the handle's scratch, so each switch of stream
    must wait for the other stream's work.  Interleaved with no explicit
    synchronisation, every product equals the one made alone."""
    import torch
    import hodlr3d_amd as H3
    x = np.random.default_rng(101).uniform(-1, 1, (60000, 3))
    h = H3.HODLR3D(4, 1.0, 8, 7, x)
    h.assemble()
    qs = [np.random.default_rng(102 + i).standard_normal(len(x)) for i in range(6)]
    ref = [h.matvec(q) for q in qs]
    torch.cuda.synchronize()
    outs = []
    for i, q in enumerate(qs):
        if i % 2:
            outs.append(h.matvec(torch.from_numpy(q).cuda()))
        else:
            outs.append(h.matvec(q))
    torch.cuda.synchronize()
    for i, (o, r) in enumerate(zip(outs, ref)):
        o = o.cpu().numpy() if hasattr(o, "cpu") else o
        assert np.array_equal(o, r), i
    h.close()


def test_allocation_failures_leave_a_consistent_handle():
    """An injected device allocation failure (h3d_debug_fail_alloc) part way
    through a copy: prepare_precision keeps nothing of the new copy and the
    old one still serves products; a move that already freed some of the old
    chunks drops the factors and refuses products instead of reading freed
    memory."""
    import hodlr3d_amd as H3
    from hodlr3d_amd import _native as nat
    x = uniform(40000, 111)
    q = np.random.default_rng(112).standard_normal(len(x))
    h = H3.HODLR3D(4, 1.0, 8, 7, x)
    h.assemble()
    y = h.matvec(q)
    nchunks_guess = 3
    for after in (0, nchunks_guess):
        nat.lib().h3d_debug_fail_alloc(after)
        try:
            with pytest.raises(H3.H3DError) as e:
                h.prepare_precision("bf16")
            assert e.value.code == nat.H3D_ERR_OOM
        finally:
            nat.lib().h3d_debug_fail_alloc(-1)
        assert h.stats()["factor_bytes"][2] == 0
        assert np.array_equal(h.matvec(q), y)
    h.prepare_precision("bf16")  # a retry without the failure builds it whole
    assert h.stats()["factor_bytes"][2] > 0
    assert np.linalg.norm(h.matvec(q, prec="bf16") - y) / np.linalg.norm(y) < 3e-2
    # a move failing after its first chunk
    nat.lib().h3d_debug_fail_alloc(1)
    try:
        with pytest.raises(H3.H3DError) as e:
            h.move_precision("fp32")
        assert e.value.code == nat.H3D_ERR_OOM
    finally:
        nat.lib().h3d_debug_fail_alloc(-1)
    with pytest.raises(H3.H3DError) as e:
        h.matvec(q)
    assert e.value.code == nat.H3D_ERR_STATE
    with pytest.raises(H3.H3DError) as e:
        h.assemble()
    assert e.value.code == nat.H3D_ERR_STATE
    h.close()
    # a failed move before any chunk moved keeps the old copy whole
    g = H3.HODLR3D(4, 1.0, 8, 7, x)
    g.assemble()
    nat.lib().h3d_debug_fail_alloc(0)
    try:
        with pytest.raises(H3.H3DError):
            g.move_precision("fp32")
    finally:
        nat.lib().h3d_debug_fail_alloc(-1)
    assert np.array_equal(g.matvec(q), y)
    g.close()
