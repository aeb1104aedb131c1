"""dh_block_dots_noalloc: the block dot products without any allocation (partial sums in a fixed
file-scope device array, result left in device memory), against dh_block_dots, through the C ABI.
Two calls in a row, the second on a smaller input, so that the second reads none of the first's
partial sums."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bind(capi):
    return capi.gpu().dh_block_dots_noalloc


@pytest.mark.parametrize("k", [1, 4])
def test_noalloc_dots_match_dots(arch, k):
    from diagham_amd import capi
    noalloc = _bind(capi)
    rng = np.random.default_rng(5 + k)
    for dim, nb in ((300001, 37), (70003, 5)):
        print("noalloc dots: k=%d dim=%d blocks=%d" % (k, dim, nb), flush=True)
        V = rng.standard_normal(nb * dim * k)
        W = rng.standard_normal(dim * k)
        dV = capi.DeviceBuffer.from_host(arch, V)
        dW = capi.DeviceBuffer.from_host(arch, W)
        dOut = capi.DeviceBuffer(arch, nb * k * k)
        ref = np.zeros(nb * k * k)
        arch.check(capi.gpu().dh_block_dots(arch.ctx, dV.ptr, nb, dW.ptr, dim, k, capi.dptr(ref)), "dh_block_dots")
        arch.check(noalloc(arch.ctx, dV.ptr, nb, dW.ptr, dim, k, dOut.ptr), "dh_block_dots_noalloc")
        arch.sync()
        out = dOut.download()
        assert np.allclose(out, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
        # numpy check of the layout: out[(b * k + p) * k + q] = sum_rows V_b[row][p] W[row][q]
        Vb = V.reshape(nb, dim, k)
        Wr = W.reshape(dim, k)
        full = np.einsum("brp,rq->bpq", Vb, Wr).reshape(-1)
        assert np.allclose(out, full, rtol=1e-10, atol=1e-10 * np.abs(full).max())
        for b in (dV, dW, dOut):
            b.free()


def test_noalloc_refuses_oversize(arch):
    from diagham_amd import capi
    noalloc = _bind(capi)
    dV = capi.DeviceBuffer(arch, 16)
    dOut = capi.DeviceBuffer(arch, 16)
    # nb * k * k = 2^20 partial sums per chunk exceed the 2^19 doubles of the scratch array: refused
    # before any launch (the basis pointer is never read)
    assert noalloc(arch.ctx, dV.ptr, 1 << 12, dV.ptr, 1, 16, dOut.ptr) == capi.DH_E_RANGE
    assert b"scratch" in capi.gpu().dh_last_error(arch.ctx)
    dV.free()
    dOut.free()
