"""Many Lz sectors in one launch: concatenated bases, persistent multi-sector product,
segmented Lanczos vector operations and the batched block Lanczos, against the oracle and the
sequential per-sector path."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

N, LZ = 6, 15
SECTORS = [0, 2, 4, 10, 24, 40, 56, 58, 60]  # twice Lz; the last ones are tiny


def coulomb(lz):
    return O.round14(O.coulomb_pseudopotentials(lz))


class SectorSet:
    def __init__(self, arch, n, lz, sectors, pseudo):
        from diagham_amd import capi
        self.capi, self.arch = capi, arch
        arr = np.array(sectors, np.int32)
        h = ctypes.c_void_p()
        arch.check(capi.gpu().dh_sector_set_create(arch.ctx, ctypes.byref(h), n, lz,
                                                   arr.ctypes.data_as(capi.P_I), len(sectors)), "sector set")
        self.h = h
        self.off = np.zeros(len(sectors) + 1, np.int64)
        capi.gpu().dh_sector_set_offsets(h, self.off.ctypes.data_as(capi.P_I64))
        self.total = int(self.off[-1])
        f, po = capi.sphere_factors(lz, pseudo)
        arch.check(capi.gpu().dh_sector_set_factors(h, capi.dptr(f), po.ctypes.data_as(capi.P_I), f.size, 0.0),
                   "factors")

    def states(self):
        out = np.zeros(self.total, np.uint64)
        self.capi.gpu().dh_sector_set_get_states(self.h, out.ctypes.data_as(self.capi.P_U64))
        return out

    def hv(self, x, k, active=None, y0=None):
        c = self.capi
        dx = c.DeviceBuffer.from_host(self.arch, x)
        dy = c.DeviceBuffer.from_host(self.arch, y0 if y0 is not None else np.zeros_like(x))
        am = None if active is None else bytes(np.asarray(active, np.uint8))
        self.arch.check(c.gpu().dh_sectors_hv(self.h, dx.ptr, dy.ptr, k, 0, am), "sectors hv")
        return dy.download()

    def close(self):
        self.capi.gpu().dh_sector_set_destroy(self.h)


@pytest.mark.parametrize("lpr,row_mode", [(1, 1), (4, 1), (1, 3), (2, 3), (1, 4), (2, 4)])
def test_sector_set_basis_and_products(arch, lpr, row_mode):
    V = coulomb(LZ)
    S = SectorSet(arch, N, LZ, SECTORS, V)
    S.capi.gpu().dh_sector_set_launch(S.h, lpr, 256)
    assert S.capi.gpu().dh_sector_set_row_mode(S.h, row_mode) == 0
    states = S.states()
    fac = O.Factors(LZ, V)
    for k in (1, 4):
        X = O.splitmix(9 + k, S.total * k)
        active = [1] * len(SECTORS)
        active[1] = 0
        y0 = O.splitmix(99, S.total * k)
        Y = S.hv(X, k, active, y0)
        for i, lz in enumerate(SECTORS):
            ref = O.FermionSphere(N, lz, LZ)
            a, b = S.off[i], S.off[i + 1]
            assert np.array_equal(states[a:b], ref.states())
            if not active[i]:
                assert np.array_equal(Y[a * k:b * k], y0[a * k:b * k])  # inactive sectors untouched
                continue
            Xs = X[a * k:b * k].reshape(-1, k)
            Ys = Y[a * k:b * k].reshape(-1, k)
            for v in range(k):
                y_ref = O.fs_hv(ref, fac, np.ascontiguousarray(Xs[:, v]))
                assert np.linalg.norm(Ys[:, v] - y_ref) <= 1e-12 * max(np.linalg.norm(y_ref), 1e-300), (lz, k, v)
    S.close()


@pytest.mark.parametrize("k,row_mode", [(1, 1), (2, 1), (8, 1), (1, 3), (4, 3), (1, 4), (2, 4)])
def test_sector_products_beyond_32_orbitals(arch, k, row_mode):
    """36 orbitals: one- and two-vector products keep 64-bit state words (the 32-bit form is for
    at most 32 orbitals) and match the oracle in every sector."""
    n, lz_max, sectors = 4, 35, [0, 2, 10, 60]
    V = coulomb(lz_max)
    S = SectorSet(arch, n, lz_max, sectors, V)
    assert S.capi.gpu().dh_sector_set_row_mode(S.h, row_mode) == 0
    fac = O.Factors(lz_max, V)
    X = O.splitmix(40 + k, S.total * k)
    Y = S.hv(X, k, [1] * len(sectors), np.zeros(S.total * k))
    for i, lz in enumerate(sectors):
        ref = O.FermionSphere(n, lz, lz_max)
        a, b = S.off[i], S.off[i + 1]
        assert b - a == ref.dim
        Xs = X[a * k:b * k].reshape(-1, k)
        Ys = Y[a * k:b * k].reshape(-1, k)
        for v in range(k):
            y_ref = O.fs_hv(ref, fac, np.ascontiguousarray(Xs[:, v]))
            assert np.linalg.norm(Ys[:, v] - y_ref) <= 1e-12 * max(np.linalg.norm(y_ref), 1e-300), (lz, k, v)
    S.close()


@pytest.mark.parametrize("n,lz_max,sectors", [(4, 31, [0, 4, 30, 80]), (4, 33, [0, 2, 12, 90])])
@pytest.mark.parametrize("k,row_mode,lpr,block", [(1, 1, 1, 1024), (2, 1, 2, 256), (1, 4, 1, 512), (8, 1, 2, 1024),
                                                  (1, 3, 4, 256)])
def test_sector_products_at_the_14_bit_rank_split(arch, n, lz_max, sectors, k, row_mode, lpr, block):
    """32 and 34 orbitals: the shared 16-bit low rank table in LDS, per-sector high tables of 2^18
    and 2^20 entries, 32-bit and 64-bit state words, against the oracle in every sector."""
    V = coulomb(lz_max)
    S = SectorSet(arch, n, lz_max, sectors, V)
    assert S.capi.gpu().dh_sector_set_row_mode(S.h, row_mode) == 0
    assert S.capi.gpu().dh_sector_set_launch(S.h, lpr, block) == 0
    states = S.states()
    fac = O.Factors(lz_max, V)
    X = O.splitmix(70 + k, S.total * k)
    Y = S.hv(X, k, [1] * len(sectors), np.zeros(S.total * k))
    for i, lz in enumerate(sectors):
        ref = O.FermionSphere(n, lz, lz_max)
        a, b = S.off[i], S.off[i + 1]
        assert b - a == ref.dim
        assert np.array_equal(states[a:b], ref.states())
        Xs = X[a * k:b * k].reshape(-1, k)
        Ys = Y[a * k:b * k].reshape(-1, k)
        for v in range(k):
            y_ref = O.fs_hv(ref, fac, Xs[:, v])
            assert np.linalg.norm(Ys[:, v] - y_ref) <= 1e-12 * np.linalg.norm(y_ref), (lz, k, v)
    S.close()


def test_sector_element_counts(arch):
    """Per-sector matrix-element counts of a set equal the oracle's count of each sector."""
    V = coulomb(LZ)
    S = SectorSet(arch, N, LZ, SECTORS, V)
    counts = np.zeros(len(SECTORS), np.int64)
    arch.check(S.capi.gpu().dh_sector_set_count_elements(S.h, counts.ctypes.data_as(S.capi.P_I64)), "count")
    fac = O.Factors(LZ, V)
    for i, lz in enumerate(SECTORS):
        assert counts[i] == O.fs_count_elements(O.FermionSphere(N, lz, LZ), fac), lz
    S.close()


# (k, basis blocks, axpy workgroups per CU, dots rows per tile): the unrolled and remainder block
# loops of the axpy, one and several LDS slabs (k = 16: 16 blocks per slab), short and long tiles
@pytest.mark.parametrize("k,nb,groups,rows", [(4, 3, 8, 0), (1, 11, 8, 0), (1, 11, 2, 256), (2, 5, 32, 0),
                                              (8, 2, 8, 512), (16, 17, 8, 0)])
def test_segmented_vector_operations(arch, k, nb, groups, rows):
    c = __import__("diagham_amd.capi", fromlist=["capi"])
    S = SectorSet(arch, N, LZ, SECTORS, coulomb(LZ))
    arch.check(c.gpu().dh_sector_set_blas_launch(S.h, groups, rows), "blas launch")
    T = S.total
    rng = np.random.default_rng(2)
    Vb = rng.standard_normal(nb * T * k)
    W = rng.standard_normal(T * k)
    dV = c.DeviceBuffer.from_host(arch, Vb)
    dW = c.DeviceBuffer.from_host(arch, W)
    active = np.ones(len(SECTORS), np.uint8)
    active[3] = 0
    am = bytes(active)
    act = [i for i in range(len(SECTORS)) if active[i]]
    out = np.zeros(len(act) * nb * k * k)
    arch.check(c.gpu().dh_sectors_block_dots(S.h, dV.ptr, nb, dW.ptr, k, am, c.dptr(out)), "dots")
    out = out.reshape(len(act), nb * k, k)
    Vm = Vb.reshape(nb, T, k)
    Wm = W.reshape(T, k)
    for a, i in enumerate(act):
        lo, hi = S.off[i], S.off[i + 1]
        ref = np.concatenate([Vm[b, lo:hi].T @ Wm[lo:hi] for b in range(nb)], axis=0)
        assert np.allclose(out[a], ref, rtol=1e-12, atol=1e-10)
    C = rng.standard_normal((len(act), nb * k, k))
    arch.check(c.gpu().dh_sectors_block_axpy(S.h, dV.ptr, nb, dW.ptr, k, am, c.dptr(np.ascontiguousarray(C))), "axpy")
    W2 = dW.download().reshape(T, k)
    for a, i in enumerate(act):
        lo, hi = S.off[i], S.off[i + 1]
        ref = Wm[lo:hi] + sum(Vm[b, lo:hi] @ C[a, b * k:(b + 1) * k] for b in range(nb))
        assert np.allclose(W2[lo:hi], ref, rtol=1e-12, atol=1e-10)
    lo, hi = S.off[3], S.off[4]
    assert np.array_equal(W2[lo:hi], Wm[lo:hi])
    M = rng.standard_normal((len(act), k, k))
    dW.upload(W)
    arch.check(c.gpu().dh_sectors_block_rmul(S.h, dW.ptr, k, am, c.dptr(np.ascontiguousarray(M))), "rmul")
    W3 = dW.download().reshape(T, k)
    for a, i in enumerate(act):
        lo, hi = S.off[i], S.off[i + 1]
        assert np.allclose(W3[lo:hi], Wm[lo:hi] @ M[a], rtol=1e-12, atol=1e-10)
    S.close()


def solve(arch, n, lz, sectors, nev, k, batched, max_iter=2000):
    from diagham_amd import capi
    V = capi.coulomb_pseudopotentials(lz, True)
    arr = np.array(sectors, np.int32)
    ev = np.zeros(len(sectors) * nev)
    prods = np.zeros(len(sectors), np.int32)
    conv = np.zeros(len(sectors), np.int32)
    secs = ctypes.c_double()
    rc = capi.host().dhh_sectors_solve(arch.h, n, lz, arr.ctypes.data_as(capi.P_I), len(sectors), capi.dptr(V), nev,
                                       k, max_iter, 1e-14, 0x5EED0000, int(batched), 8.0, capi.dptr(ev),
                                       prods.ctypes.data_as(capi.P_I), conv.ctypes.data_as(capi.P_I),
                                       ctypes.byref(secs))
    assert rc == 0, capi.host().dhh_last_error().decode()
    return ev.reshape(len(sectors), nev), prods, conv, secs.value


def test_batched_equals_sequential_and_oracle(arch):
    sectors = list(range(0, 61, 2))  # every Lz >= 0 sector of N = 6, 2S = 15
    nev = 5
    eb, pb, cb, tb = solve(arch, N, LZ, sectors, nev, 4, True)
    es, ps, cs, ts = solve(arch, N, LZ, sectors, nev, 1, False)
    assert cb.all() and cs.all()
    for i, lz in enumerate(sectors):
        d = O.lib().or_fs_dimension(N, LZ, lz)
        m = min(nev, d)
        np.testing.assert_allclose(eb[i, :m], es[i, :m], rtol=0, atol=1e-10)
        if d <= 400:  # oracle dense diagonalisation
            ref = O.FermionSphere(N, lz, LZ)
            fac = O.Factors(LZ, coulomb(LZ))
            H = np.zeros((d, d))
            for c in range(d):
                e = np.zeros(d)
                e[c] = 1.0
                H[:, c] = O.fs_hv(ref, fac, e)
            np.testing.assert_allclose(eb[i, :m], np.linalg.eigvalsh(H)[:m], rtol=0, atol=1e-10)


def test_batched_reference_sector(arch, golden):
    """the 2S = 21, N = 8 Lz = 0 sector solved inside a batch of sectors reproduces the reference"""
    ref = np.array(golden["fermions_sphere_n8_2s21_lz0"]["eigenvalues"])
    ev, prods, conv, secs = solve(arch, 8, 21, [0, 2, 4, 6], 10, 8, True)
    assert conv.all()
    np.testing.assert_allclose(ev[0], ref, rtol=0, atol=1e-10)


def test_batched_solve_grows_krylov_storage(arch, golden):
    """The batched run for the 40 lowest eigenvalues of the Lz = 0 and 1 sectors of N=10 2S=27 takes
    more one-vector steps than the initial Krylov storage holds (256 vectors); the storage grows
    within the device's free memory and the Lz = 0 sector still gives the reference eigenvalues."""
    import ctypes
    from diagham_amd import capi
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    arch.check(capi.gpu().dh_memory_info(arch.ctx, ctypes.byref(free), ctypes.byref(total)), "memory info")
    assert 0 < free.value <= total.value
    pp = capi.coulomb_pseudopotentials(27, True)
    r = capi.sectors_solve(arch, 10, 27, [0, 2], pp, 40, 1, batched=True)
    assert r["converged"].all()
    assert r["products"].max() > 256, r["products"]
    ref = np.array(golden["fermions_sphere_n10_2s27_lz0"]["eigenvalues"])
    np.testing.assert_allclose(r["eigenvalues"][0][:ref.size], ref, rtol=0, atol=1e-10)


def test_sector_autotune_picks_a_shape_that_keeps_the_eigenvalues(arch):
    """The autotuner times every launch shape per block size on calibration sectors, then their
    solve; the winner, applied to the whole set, gives the eigenvalues of the built-in launch."""
    from diagham_amd import capi
    n, lz = 8, 21
    V = capi.coulomb_pseudopotentials(lz, True)
    lzs = capi.sphere_sectors(n, lz)
    rows, best = capi.sectors_autotune(arch, n, lz, lzs, V, 4, ks=(1, 4), fraction=0.25)
    assert [r["k"] for r in rows] == [1.0, 4.0]
    for r in rows:
        assert r["lanes_per_row"] in (1, 2, 4) and r["block_size"] in (256, 1024)
        assert r["row_mode"] in (capi.INDEX_RANK, capi.INDEX_RANK_FREE, capi.INDEX_RANK_RANGE)
        assert r["product_ms"] > 0 and r["calibration_seconds"] > 0 and r["calibration_products"] > 0
    assert best["k"] in (1, 4)
    ref = capi.sectors_solve(arch, n, lz, lzs, V, 4, best["k"])
    capi.sectors_launch(best["lanes_per_row"], best["block_size"], best["row_mode"])
    try:
        tuned = capi.sectors_solve(arch, n, lz, lzs, V, 4, best["k"])
    finally:
        capi.sectors_launch(0)
    assert np.all(tuned["converged"])
    assert np.max(np.abs(tuned["eigenvalues"] - ref["eigenvalues"])) < 1e-10


@pytest.mark.parametrize("k", [1, 4])
def test_large_sectors_run_alone_in_a_batched_solve(arch, golden, k):
    """dhh_sectors_alone_rows: with the threshold at 5,000 states, the eight sectors of N=8 2S=21
    (2Lz = 0, 4, .., 80) above it go through the single-space solver one at a time and the other
    thirteen through the waves; every sector's eigenvalues equal those of the all-batched solve and
    of the reference loop to 1e-10, and Lz = 0 reproduces the reference's printed values."""
    from diagham_amd import capi
    sectors = list(range(0, 81, 4))
    assert capi.host().dhh_sectors_alone_rows(-1) == capi.DH_E_ARG
    try:
        assert capi.host().dhh_sectors_alone_rows(5000) == 0
        ea, pa, ca, _ = solve(arch, 8, 21, sectors, 10, k, True)
        assert capi.host().dhh_sectors_alone_rows(0) == 0
        eb, pb, cb, _ = solve(arch, 8, 21, sectors, 10, k, True)
    finally:
        capi.host().dhh_sectors_alone_rows(2000000)
    es, ps, cs, _ = solve(arch, 8, 21, sectors, 10, 1, False)
    assert ca.all() and cb.all() and cs.all()
    dims = [O.lib().or_fs_dimension(8, 21, lz) for lz in sectors]
    assert sum(d >= 5000 for d in dims) == 8 and sum(d < 5000 for d in dims) == 13
    for i, d in enumerate(dims):
        m = min(10, d)
        np.testing.assert_allclose(ea[i, :m], eb[i, :m], rtol=0, atol=1e-10)
        np.testing.assert_allclose(ea[i, :m], es[i, :m], rtol=0, atol=1e-10)
    ref = np.array(golden["fermions_sphere_n8_2s21_lz0"]["eigenvalues"])
    np.testing.assert_allclose(ea[0], ref, rtol=0, atol=1e-10)
