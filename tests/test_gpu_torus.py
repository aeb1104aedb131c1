"""Fermions on the torus (one Ky sector each) on the GPU: FermionOnTorus as a momentum sector set
with Nx = 1, the torus two-body factors (ParticleOnTorusGenericHamiltonian,
ParticleOnTorusCoulombHamiltonian) through the complex gather product, and the batched Lanczos
over Ky sectors, against the oracle and the exact Laughlin zero modes of V1."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


class TorusSet:
    def __init__(self, arch, n, nflux, kys, ratio=1.0, pseudo=(), level=-1):
        from diagham_amd import capi
        self.capi, self.arch = capi, arch
        kx = np.zeros(len(kys), np.int32)
        ky = np.array(kys, np.int32)
        h = ctypes.c_void_p()
        arch.check(capi.gpu().dh_momentum_sector_set_create(arch.ctx, ctypes.byref(h), n, 1, nflux,
                                                            kx.ctypes.data_as(capi.P_I), ky.ctypes.data_as(capi.P_I),
                                                            len(kys)), "torus set")
        self.h = h
        self.off = np.zeros(len(kys) + 1, np.int64)
        capi.gpu().dh_sector_set_offsets(h, self.off.ctypes.data_as(capi.P_I64))
        self.total = int(self.off[-1])
        cnt, pairs, f = capi.torus_factors(nflux, ratio, pseudo, level)
        arch.check(capi.gpu().dh_sector_set_lattice_factors(h, nflux, cnt.ctypes.data_as(capi.P_I),
                                                            pairs.ctypes.data_as(capi.P_I), capi.dptr(f), 0.0),
                   "torus factors")

    def states(self):
        out = np.zeros(self.total, np.uint64)
        self.capi.gpu().dh_sector_set_get_states(self.h, out.ctypes.data_as(self.capi.P_U64))
        return out

    def hv(self, X, k):
        c = self.capi
        xr = np.empty(self.total * k * 2)
        xr[0::2] = X.reshape(-1).real
        xr[1::2] = X.reshape(-1).imag
        dx = c.DeviceBuffer.from_host(self.arch, xr)
        dy = c.DeviceBuffer(self.arch, xr.size)
        self.arch.check(c.gpu().dh_sectors_hv(self.h, dx.ptr, dy.ptr, k, 0, None), "torus hv")
        y = dy.download()
        dx.free()
        dy.free()
        return (y[0::2] + 1j * y[1::2]).reshape(self.total, k)

    def close(self):
        self.capi.gpu().dh_sector_set_destroy(self.h)


@pytest.mark.parametrize("n,nflux", [(4, 12), (6, 18), (5, 11)])
def test_torus_basis_bit_exact(arch, n, nflux):
    kys = list(range(nflux))
    S = TorusSet(arch, n, nflux, kys, pseudo=[0.0, 1.0])
    st = S.states()
    for i, ky in enumerate(kys):
        assert np.array_equal(st[S.off[i]:S.off[i + 1]], O.torus_states_brute_force(n, nflux, ky))
    S.close()


@pytest.mark.parametrize("level,lpr", [(-1, 1), (0, 4), (1, 1)])
def test_torus_products_match_oracle(arch, level, lpr):
    n, nflux, ratio = 5, 15, 0.7
    pseudo = [0.0, 1.0, 0.0, 0.35] if level < 0 else []
    kys = [0, 3, 7, 14]
    S = TorusSet(arch, n, nflux, kys, ratio, pseudo, level)
    S.capi.gpu().dh_sector_set_launch(S.h, lpr, 256)
    fac = O.TorusFactors(nflux, ratio, pseudo=pseudo, coulomb_level=level)
    rng = np.random.default_rng(11)
    for k in (1, 4):
        X = rng.standard_normal((S.total, k)) + 1j * rng.standard_normal((S.total, k))
        Y = S.hv(X, k)
        for i, ky in enumerate(kys):
            sp = O.torus_space(n, nflux, ky)
            a, b = S.off[i], S.off[i + 1]
            for v in range(k):
                ref = O.fm_hv(sp, fac, X[a:b, v])
                assert np.linalg.norm(Y[a:b, v] - ref) <= 1e-12 * np.linalg.norm(ref)
    S.close()


def test_torus_laughlin_zero_modes_on_gpu(arch):
    """V1 at filling 1/3, N = 6, Nphi = 18 (about 1,030 states per Ky): exactly three zero modes,
    one in each of Ky0, Ky0 + 6, Ky0 + 12, and a gap above."""
    from diagham_amd import capi
    n, nflux = 6, 18
    r = capi.torus_sectors_solve(arch, n, nflux, list(range(nflux)), 3, pseudo=[0.0, 1.0], block_size=1)
    assert r["converged"].all()
    ev = r["eigenvalues"]
    zero = [ky for ky in range(nflux) if ev[ky, 0] < 1e-10]
    assert len(zero) == 3 and (zero[1] - zero[0]) % n == 0 and (zero[2] - zero[1]) % n == 0, ev[:, 0]
    assert np.abs(ev[zero, 0]).max() < 1e-10
    others = np.delete(ev[:, 0], zero)
    assert others.min() > 1e-3
    assert min(ev[zero, 1]) > 1e-3


@pytest.mark.parametrize("k", [1, 4])
def test_torus_coulomb_sectors_match_dense_oracle(arch, k):
    from diagham_amd import capi
    n, nflux, ratio = 5, 15, 0.9
    kys = [0, 1, 5, 10]
    r = capi.torus_sectors_solve(arch, n, nflux, kys, 4, ratio=ratio, coulomb_level=0, block_size=k)
    assert r["converged"].all()
    fac = O.TorusFactors(nflux, ratio, coulomb_level=0)
    for i, ky in enumerate(kys):
        H = O.dense_fm(O.torus_space(n, nflux, ky), fac)
        full = np.linalg.eigvalsh(H.real)
        if k == 1:
            # one vector per block finds each eigenvalue once, also a degenerate one (the Kx
            # pairs of a Ky sector): the lowest matches, the others are in the spectrum
            assert abs(r["eigenvalues"][i][0] - full[0]) < 1e-10
            assert all(np.abs(full - e).min() < 1e-10 for e in r["eigenvalues"][i])
        else:
            assert np.abs(r["eigenvalues"][i] - full[:4]).max() < 1e-10
    # Ky = 0, 5, 10 are related by the magnetic translation (Ky -> Ky + N): equal spectra
    assert np.abs(r["eigenvalues"][0] - r["eigenvalues"][2]).max() < 1e-10
    assert np.abs(r["eigenvalues"][0] - r["eigenvalues"][3]).max() < 1e-10


class TranslationSet(TorusSet):
    def __init__(self, arch, n, nflux, sectors, ratio=1.0, pseudo=(), level=-1):
        from diagham_amd import capi
        self.capi, self.arch = capi, arch
        kx = np.array([s[0] for s in sectors], np.int32)
        ky = np.array([s[1] for s in sectors], np.int32)
        h = ctypes.c_void_p()
        arch.check(capi.gpu().dh_torus_translation_sector_set_create(arch.ctx, ctypes.byref(h), n, nflux,
                                                                     kx.ctypes.data_as(capi.P_I),
                                                                     ky.ctypes.data_as(capi.P_I), len(sectors)),
                   "torus translation set")
        self.h = h
        self.off = np.zeros(len(sectors) + 1, np.int64)
        capi.gpu().dh_sector_set_offsets(h, self.off.ctypes.data_as(capi.P_I64))
        self.total = int(self.off[-1])
        cnt, pairs, f = capi.torus_factors(nflux, ratio, pseudo, level)
        arch.check(capi.gpu().dh_sector_set_lattice_factors(h, nflux, cnt.ctypes.data_as(capi.P_I),
                                                            pairs.ctypes.data_as(capi.P_I), capi.dptr(f), 0.0),
                   "torus factors")

    def orbit_data(self):
        orb = np.zeros(self.total, np.uint8)
        ro = np.zeros(self.total, np.uint64)
        self.arch.check(self.capi.gpu().dh_sector_set_torus_data(self.h, orb.ctypes.data,
                                                                 ro.ctypes.data_as(self.capi.P_U64)), "torus data")
        return orb, ro


def all_sectors(n, nflux):
    from math import gcd
    g = gcd(n, nflux)
    return [(kx, ky) for ky in range(nflux) for kx in range(g)]


@pytest.mark.parametrize("n,nflux", [(4, 12), (6, 12), (5, 15), (3, 9), (4, 10)])
def test_translation_basis_bit_exact(arch, n, nflux):
    """Representatives, orbit sizes and reordering-sign bits of every (Kx, Ky) sector against the
    oracle (whose sectors are pinned by splitting each Ky spectrum exactly)."""
    sectors = [s for s in all_sectors(n, nflux) if O.TorusTranslations(n, nflux, *s).dim > 0]
    S = TranslationSet(arch, n, nflux, sectors, pseudo=[0.0, 1.0])
    st = S.states()
    orb, ro = S.orbit_data()
    for i, (kx, ky) in enumerate(sectors):
        ref = O.TorusTranslations(n, nflux, kx, ky)
        a, b = S.off[i], S.off[i + 1]
        assert np.array_equal(st[a:b], ref.states())
        assert np.array_equal(orb[a:b].astype(np.int32), ref.orbits())
        assert np.array_equal(ro[a:b], ref.reorder())
    S.close()


@pytest.mark.parametrize("n,nflux,level,lpr", [(4, 12, -1, 1), (6, 12, 0, 4), (5, 15, 1, 1), (4, 10, -1, 2)])
def test_translation_products_match_oracle(arch, n, nflux, level, lpr):
    ratio = 0.85
    pseudo = [0.1, 1.0, 0.0, 0.3] if level < 0 else []
    sectors = [s for s in all_sectors(n, nflux) if O.TorusTranslations(n, nflux, *s).dim > 0][:12]
    S = TranslationSet(arch, n, nflux, sectors, ratio, pseudo, level)
    S.capi.gpu().dh_sector_set_launch(S.h, lpr, 256)
    fac = O.TorusFactors(nflux, ratio, pseudo=pseudo, coulomb_level=level)
    rng = np.random.default_rng(7)
    for k in (1, 4):
        X = rng.standard_normal((S.total, k)) + 1j * rng.standard_normal((S.total, k))
        Y = S.hv(X, k)
        for i, (kx, ky) in enumerate(sectors):
            sp = O.TorusTranslations(n, nflux, kx, ky)
            a, b = S.off[i], S.off[i + 1]
            for v in range(k):
                ref = sp.hv(fac, X[a:b, v])
                assert np.linalg.norm(Y[a:b, v] - ref) <= 1e-12 * max(np.linalg.norm(ref), 1e-300), (kx, ky, v)
    S.close()


@pytest.mark.parametrize("k", [1, 4])
def test_translation_sectors_reproduce_ky_sectors(arch, k):
    """N = 6 at Nphi = 18 (Coulomb, lowest Landau level): the lowest eigenvalue of each Ky sector
    solved without translations equals the lowest over Kx of the (Kx, Ky) sectors, and every
    translation sector matches the dense oracle."""
    from diagham_amd import capi
    n, nflux = 6, 18
    kys = [0, 1, 3]
    r0 = capi.torus_sectors_solve(arch, n, nflux, kys, 2, coulomb_level=0, block_size=k)
    assert r0["converged"].all()
    sec = [(kx, ky) for ky in kys for kx in range(6)]
    r = capi.torus_translation_sectors_solve(arch, n, nflux, [s[0] for s in sec], [s[1] for s in sec], 2,
                                             coulomb_level=0, block_size=k)
    assert r["converged"].all()
    ev = r["eigenvalues"]
    for j, ky in enumerate(kys):
        lows = np.sort(ev[[i for i, s in enumerate(sec) if s[1] == ky]].reshape(-1))
        assert abs(lows[0] - r0["eigenvalues"][j, 0]) < 1e-10
    fac = O.TorusFactors(nflux, 1.0, coulomb_level=0)
    for i, (kx, ky) in enumerate(sec[:6]):
        full = np.linalg.eigvalsh(O.TorusTranslations(n, nflux, kx, ky).dense(fac))
        assert abs(ev[i][0] - full[0]) < 1e-10
        assert all(np.abs(full - e).min() < 1e-10 for e in ev[i])
        if k > 1:
            assert np.abs(ev[i] - full[:2]).max() < 1e-10
