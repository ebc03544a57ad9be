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
