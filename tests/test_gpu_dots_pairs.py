"""k = 1 block dot products read two rows per lane when the basis and W are 16-byte aligned and the
block stride is even, with an odd first and last row of a range on their own lanes; otherwise one
row per lane.  Both forms against numpy, through the C ABI: single space (dh_block_dots, and the
allocation-free form that shares its kernel) and sector sets (dh_sectors_block_dots) with odd
sector offsets and odd tile lengths."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim", [1, 2, 3, 1025, 4103, 200000, 200001])
@pytest.mark.parametrize("nb", [1, 5, 8, 9, 17])
def test_single_space_dots(arch, dim, nb):
    from diagham_amd import capi
    rng = np.random.default_rng(dim * 31 + nb)
    V = rng.standard_normal(nb * dim)
    w = rng.standard_normal(dim)
    dV = capi.DeviceBuffer.from_host(arch, V)
    dW = capi.DeviceBuffer.from_host(arch, w)
    out = np.zeros(nb)
    arch.check(capi.gpu().dh_block_dots(arch.ctx, dV.ptr, nb, dW.ptr, dim, 1, capi.dptr(out)), "dh_block_dots")
    ref = V.reshape(nb, dim) @ w
    assert np.allclose(out, ref, rtol=1e-12, atol=1e-12 * np.sqrt(dim)), (dim, nb)
    dV.free()
    dW.free()


@pytest.mark.parametrize("shift_v,shift_w", [(1, 0), (0, 1), (1, 1)])
def test_single_space_dots_unaligned(arch, shift_v, shift_w):
    """V or W starting 8 bytes into an allocation (not 16-byte aligned): the one-row form."""
    from diagham_amd import capi
    dim, nb = 100002, 9
    rng = np.random.default_rng(5)
    V = rng.standard_normal(nb * dim + 1)
    w = rng.standard_normal(dim + 1)
    dV = capi.DeviceBuffer.from_host(arch, V)
    dW = capi.DeviceBuffer.from_host(arch, w)
    out = np.zeros(nb)
    pv, pw = dV.ptr + 8 * shift_v, dW.ptr + 8 * shift_w
    arch.check(capi.gpu().dh_block_dots(arch.ctx, pv, nb, pw, dim, 1, capi.dptr(out)), "dh_block_dots")
    ref = V[shift_v:shift_v + nb * dim].reshape(nb, dim) @ w[shift_w:shift_w + dim]
    assert np.allclose(out, ref, rtol=1e-12, atol=1e-12 * np.sqrt(dim))
    dV.free()
    dW.free()


@pytest.mark.parametrize("rows_per_tile", [0, 256, 257])
def test_sector_dots_odd_offsets(arch, rows_per_tile):
    """Sectors of N=6 2S=15 with 2Lz = 0 .. 58: an even total (1488 rows) and sectors starting at
    odd rows (1293, 1485, 1487); tiles of the default length, of 256 rows, and of 257 rows (tiles
    starting at odd rows inside a sector)."""
    from diagham_amd import capi
    n, lz_max, sectors = 6, 15, [0, 2, 4, 10, 24, 40, 56, 58]
    S = capi.SphereSectorSet(arch, n, lz_max, sectors, O.round14(O.coulomb_pseudopotentials(lz_max)))
    assert S.total == 1488 and S.total % 2 == 0 and any(o % 2 for o in S.offsets[:-1])
    arch.check(capi.gpu().dh_sector_set_blas_launch(S.h, 16, rows_per_tile), "blas launch")
    nb = 11
    rng = np.random.default_rng(rows_per_tile + 3)
    V = rng.standard_normal(nb * S.total)
    w = rng.standard_normal(S.total)
    dV = capi.DeviceBuffer.from_host(arch, V)
    dW = capi.DeviceBuffer.from_host(arch, w)
    active = np.ones(len(sectors), np.uint8)
    active[1] = 0
    act = [i for i in range(len(sectors)) if active[i]]
    out = np.zeros(len(act) * nb)
    arch.check(capi.gpu().dh_sectors_block_dots(S.h, dV.ptr, nb, dW.ptr, 1, bytes(active), capi.dptr(out)), "dots")
    Vm = V.reshape(nb, S.total)
    for a, i in enumerate(act):
        lo, hi = S.offsets[i], S.offsets[i + 1]
        assert np.allclose(out[a * nb:(a + 1) * nb], Vm[:, lo:hi] @ w[lo:hi], rtol=1e-12, atol=1e-12), (i, lo, hi)
    S.close()
