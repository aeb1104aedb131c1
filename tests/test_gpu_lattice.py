"""Chern-band lattice (complex FP64): momentum sector sets, complex product and the complex
block Lanczos over many sectors, against the oracle."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

T1, T2, TPP = 1.0, 0.5, -0.5


def host_factors(capi, nx, ny):
    cnt = np.zeros(nx * ny, np.int32)
    npairs = nx * ny * (nx * ny - 1) // 2
    pairs = np.zeros(2 * npairs, np.int32)
    nf = ctypes.c_int64(npairs * npairs)
    npr = ctypes.c_int(npairs)
    f = np.zeros(2 * npairs * npairs)
    rc = capi.host().dhh_checkerboard_factors(nx, ny, T1, T2, TPP, 0.0, 1.0, 0.0, 1, cnt.ctypes.data_as(capi.P_I),
                                              pairs.ctypes.data_as(capi.P_I), ctypes.byref(npr), capi.dptr(f),
                                              ctypes.byref(nf))
    assert rc == 0
    return cnt, pairs[:2 * npr.value], f[:2 * nf.value]


class MomentumSet:
    def __init__(self, arch, n, nx, ny, sectors):
        from diagham_amd import capi
        self.capi, self.arch = capi, arch
        kx = np.array([s[0] for s in sectors], np.int32)
        ky = np.array([s[1] for s in sectors], np.int32)
        h = ctypes.c_void_p()
        arch.check(capi.gpu().dh_momentum_sector_set_create(arch.ctx, ctypes.byref(h), n, nx, ny,
                                                            kx.ctypes.data_as(capi.P_I), ky.ctypes.data_as(capi.P_I),
                                                            len(sectors)), "momentum set")
        self.h = h
        self.off = np.zeros(len(sectors) + 1, np.int64)
        capi.gpu().dh_sector_set_offsets(h, self.off.ctypes.data_as(capi.P_I64))
        self.total = int(self.off[-1])
        cnt, pairs, f = host_factors(capi, nx, ny)
        arch.check(capi.gpu().dh_sector_set_lattice_factors(h, nx * ny, cnt.ctypes.data_as(capi.P_I),
                                                            pairs.ctypes.data_as(capi.P_I), capi.dptr(f), 0.0),
                   "lattice factors")

    def states(self):
        out = np.zeros(self.total, np.uint64)
        self.capi.gpu().dh_sector_set_get_states(self.h, out.ctypes.data_as(self.capi.P_U64))
        return out

    def hv(self, X, k):
        """X: complex (total, k) -> H X"""
        c = self.capi
        xr = np.empty(self.total * k * 2)
        xr[0::2] = X.reshape(-1).real
        xr[1::2] = X.reshape(-1).imag
        dx = c.DeviceBuffer.from_host(self.arch, xr)
        dy = c.DeviceBuffer(self.arch, xr.size)
        self.arch.check(c.gpu().dh_sectors_hv(self.h, dx.ptr, dy.ptr, k, 0, None), "lattice hv")
        y = dy.download()
        return (y[0::2] + 1j * y[1::2]).reshape(self.total, k)

    def close(self):
        self.capi.gpu().dh_sector_set_destroy(self.h)


@pytest.mark.parametrize("lpr,row_mode,block", [(1, 1, 256), (2, 1, 512), (4, 1, 256), (1, 3, 256), (2, 3, 512),
                                                (2, 1, 1024), (1, 4, 256)])
def test_momentum_basis_and_complex_products(arch, lpr, row_mode, block):
    n, nx, ny = 4, 4, 3
    sectors = [(0, 0), (1, 2), (3, 1)]
    S = MomentumSet(arch, n, nx, ny, sectors)
    S.capi.gpu().dh_sector_set_launch(S.h, lpr, block)
    assert S.capi.gpu().dh_sector_set_row_mode(S.h, row_mode) == 0
    states = S.states()
    u, _ = O.checkerboard_lower_band(nx, ny, T1, T2, TPP)
    fac = O.LatticeFactors(nx, ny, u)
    rng = np.random.default_rng(3)
    for k in (1, 4):
        X = rng.standard_normal((S.total, k)) + 1j * rng.standard_normal((S.total, k))
        Y = S.hv(X, k)
        for i, (kx, ky) in enumerate(sectors):
            ref = O.MomentumFermions(n, nx, ny, kx, ky)
            a, b = S.off[i], S.off[i + 1]
            assert np.array_equal(states[a:b], ref.states())
            for v in range(k):
                y_ref = O.fm_hv(ref, fac, X[a:b, v])
                assert np.linalg.norm(Y[a:b, v] - y_ref) <= 1e-12 * np.linalg.norm(y_ref), (kx, ky, k, v)
    S.close()


def test_complex_block_lanczos_matches_dense(arch):
    from diagham_amd import capi
    n, nx, ny = 4, 4, 3
    sectors = [(0, 0), (1, 2), (2, 1)]
    kx = np.array([s[0] for s in sectors], np.int32)
    ky = np.array([s[1] for s in sectors], np.int32)
    nev = 5
    u, _ = O.checkerboard_lower_band(nx, ny, T1, T2, TPP)
    fac = O.LatticeFactors(nx, ny, u)
    for k in (1, 4):
        ev = np.zeros(len(sectors) * nev)
        prods = np.zeros(len(sectors), np.int32)
        conv = np.zeros(len(sectors), np.int32)
        secs = ctypes.c_double()
        rc = capi.host().dhh_lattice_sectors_solve(arch.h, n, nx, ny, T1, T2, TPP, 0.0, 1.0, 0.0, 1,
                                                   kx.ctypes.data_as(capi.P_I), ky.ctypes.data_as(capi.P_I),
                                                   len(sectors), nev, k, 2000, 1e-14, 7, 4.0, capi.dptr(ev),
                                                   prods.ctypes.data_as(capi.P_I), conv.ctypes.data_as(capi.P_I),
                                                   ctypes.byref(secs))
        assert rc == 0, capi.host().dhh_last_error().decode()
        ev = ev.reshape(len(sectors), nev)
        for i, (sx, sy) in enumerate(sectors):
            ref = O.MomentumFermions(n, nx, ny, sx, sy)
            D = ref.dim
            H = np.zeros((D, D), complex)
            for c in range(D):
                e = np.zeros(D, complex)
                e[c] = 1.0
                H[:, c] = O.fm_hv(ref, fac, e)
            np.testing.assert_allclose(ev[i], np.linalg.eigvalsh(H)[:nev], rtol=0, atol=1e-10)


def test_momentum_element_counts_match_oracle_rows(arch):
    """Matrix elements per sector of a momentum set: the terms the oracle lists row by row."""
    from diagham_amd import capi
    n, nx, ny = 4, 4, 3
    sectors = [(0, 0), (1, 2), (3, 1)]
    S = capi.MomentumSectorSet(arch, n, nx, ny, [s[0] for s in sectors], [s[1] for s in sectors])
    counts = S.count_elements()
    u, _ = O.checkerboard_lower_band(nx, ny, T1, T2, TPP)
    fac = O.LatticeFactors(nx, ny, u)
    for i, (kx, ky) in enumerate(sectors):
        ref = O.MomentumFermions(n, nx, ny, kx, ky)
        expect = sum(len(O.fm_row_elements(ref, fac, r)[0]) for r in range(ref.dim))
        assert counts[i] == expect, (kx, ky, counts[i], expect)
    S.close()
