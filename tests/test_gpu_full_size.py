"""GPU parity at the full sizes of the measured workloads, against the oracle.

* Heisenberg chain L=28, Sz=0 (40,116,600 states): basis bit-exact, the whole one-vector product
  against the oracle's scatter loop (SpinChainHamiltonian.cc:257), symmetry x.Hy = y.Hx.
* Fermions on the sphere N=12, 2S=33, all 133 Lz >= 0 sectors in one set (278,059,216 rows): the
  batched k=8 product of the whole set with the launch shape of the batched solves, checked per
  sector on row slices covering at least 10 % of each checked sector (gather form of the oracle,
  AbstractQHEOnSphereHamiltonian.cc:431-455); bases of four sectors bit-exact, Lz=0 included.
* Checkerboard Chern band, N=10 on 6x5 cells, all 30 (kx, ky) sectors (30,045,015 rows): the
  complex k=4 product (ParticleOnLatticeTimeReversalBreakingSingleBandHamiltonian.h:370) with the
  large-set launch shape, checked on slices of two sectors; their bases bit-exact.

Vectors stay on the device; only the checked slices come back to the host.
"""
import os

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def download(arch, buf, first, count):
    """count doubles of a DeviceBuffer from double index `first`"""
    from diagham_amd import capi
    out = np.empty(count)
    arch.check(capi.gpu().dh_memcpy_d2h(arch.ctx, out.ctypes.data, buf.ptr + 8 * first, 8 * count), "d2h slice")
    return out


def slices(dim, fraction=0.1, pieces=3):
    """`pieces` equal row windows, spread from the start to the end of [0, dim), covering at least
    `fraction` of it (all of it for small sectors)"""
    if dim <= 200000:
        return [(0, dim)]
    w = int(np.ceil(dim * fraction / pieces))
    return [(a, a + w) for a in np.linspace(0, dim - w, pieces).astype(np.int64).tolist()]


def test_heisenberg_l28_full_product(arch):
    from diagham_amd import capi
    L = 28
    print("L=28 Sz=0: basis, product, oracle scatter over all rows", flush=True)
    p = capi.Problem.spin_chain(arch, L, 0)
    ref = O.SpinChain(L, 0)
    assert p.dim == ref.dim == 40116600
    assert np.array_equal(p.states(), ref.states())
    x = O.splitmix(0x5EED0000, p.dim)
    y = p.hv(x)
    J = np.ones(L)
    y_ref = np.zeros_like(x)
    O.lib().or_spin_hv_rows(ref.h, O.dptr(J), O.dptr(J), 1, O.dptr(x), O.dptr(y_ref), 0, p.dim, THREADS)
    err = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    assert err < 1e-12, err
    x2 = O.splitmix(17, p.dim)
    y2 = p.hv(x2)
    assert abs(x2 @ y - x @ y2) <= 1e-12 * np.linalg.norm(x) * np.linalg.norm(y2)
    p.close()


def test_sphere_n12_2s33_all_sectors_k8(arch):
    from diagham_amd import capi
    n, lz_max, k = 12, 33, 8
    V = O.round14(O.coulomb_pseudopotentials(lz_max))
    lz = capi.sphere_sectors(n, lz_max)
    assert len(lz) == 133
    print("N=12 2S=33: set of 133 sectors, batched k=8 product", flush=True)
    S = capi.SphereSectorSet(arch, n, lz_max, lz, V)
    assert S.total == 278059216
    S.set_launch(2, 1024)
    dx = capi.DeviceBuffer(arch, S.total * k)
    dy = capi.DeviceBuffer(arch, S.total * k)
    arch.check(capi.gpu().dh_fill_random(arch.ctx, dx.ptr, S.total * k, 0x5EED0000), "fill")
    S.hv_device(dx.ptr, dy.ptr, k)
    arch.sync()
    states = np.zeros(S.total, np.uint64)
    arch.check(capi.gpu().dh_sector_set_get_states(S.h, states.ctypes.data_as(capi.P_U64)), "states")
    fac = O.Factors(lz_max, V)
    for tlz in (0, 2, 120, 264):
        i = lz.index(tlz)
        a, b = int(S.offsets[i]), int(S.offsets[i + 1])
        print("  sector 2Lz=%d (%d states): basis, oracle rows" % (tlz, b - a), flush=True)
        ref = O.FermionSphere(n, tlz, lz_max)
        assert ref.dim == b - a
        assert np.array_equal(states[a:b], ref.states())
        X = download(arch, dx, a * k, (b - a) * k)
        for r0, r1 in slices(b - a):
            Y = download(arch, dy, (a + r0) * k, (r1 - r0) * k)
            Yr = O.fs_hv_rows(ref, fac, X, r0, r1, k=k, threads=THREADS)
            for v in range(k):
                yv, rv = Y[v::k], Yr[v::k]
                assert np.linalg.norm(yv - rv) <= 1e-12 * np.linalg.norm(rv), (tlz, r0, r1, v)
        del ref, X
    dx.free()
    dy.free()
    S.close()


def test_sphere_n12_2s33_lz0_k1_default_launch(arch):
    """The one-vector product of the largest sector of configuration 3 with the Hamiltonian's own
    launch choice (free-pair rows, 16-bit low rank table in LDS, one lane per row, 1024-thread
    workgroups), checked on row windows covering 10 % of the 7,764,392 rows."""
    from diagham_amd import capi
    n, lz_max = 12, 33
    V = O.round14(O.coulomb_pseudopotentials(lz_max))
    print("N=12 2S=33 Lz=0: one-vector product with the automatic launch", flush=True)
    p = capi.Problem.sphere(arch, n, 0, lz_max, V)
    assert p.dim == 7764392
    dx = capi.DeviceBuffer(arch, p.dim)
    dy = capi.DeviceBuffer(arch, p.dim)
    arch.check(capi.gpu().dh_fill_random(arch.ctx, dx.ptr, p.dim, 0x5EED0009), "fill")
    p.hv_device(dx.ptr, dy.ptr, 1)
    arch.sync()
    ref = O.FermionSphere(n, 0, lz_max)
    fac = O.Factors(lz_max, V)
    X = dx.download()
    for r0, r1 in slices(p.dim):
        Y = download(arch, dy, r0, r1 - r0)
        Yr = O.fs_hv_rows(ref, fac, X, r0, r1, k=1, threads=THREADS)
        assert np.linalg.norm(Y - Yr) <= 1e-12 * np.linalg.norm(Yr), (r0, r1)
    dx.free()
    dy.free()
    p.close()


def test_checkerboard_6x5_all_sectors_k4(arch):
    from diagham_amd import capi
    from test_gpu_lattice import MomentumSet, T1, T2, TPP
    n, nx, ny, k = 10, 6, 5, 4
    sectors = [(kx, ky) for kx in range(nx) for ky in range(ny)]
    print("N=10 on 6x5: set of 30 sectors, complex k=4 product", flush=True)
    M = MomentumSet(arch, n, nx, ny, sectors)
    assert M.total == 30045015
    arch.check(capi.gpu().dh_sector_set_launch(M.h, 2, 512), "launch")
    nd = 2 * M.total * k
    dx = capi.DeviceBuffer(arch, nd)
    dy = capi.DeviceBuffer(arch, nd)
    arch.check(capi.gpu().dh_fill_random(arch.ctx, dx.ptr, nd, 0x5EED0005), "fill")
    arch.check(capi.gpu().dh_sectors_hv(M.h, dx.ptr, dy.ptr, k, 0, None), "lattice hv")
    arch.sync()
    states = M.states()
    u, _ = O.checkerboard_lower_band(nx, ny, T1, T2, TPP)
    fac = O.LatticeFactors(nx, ny, u)
    for sec in ((0, 0), (3, 2)):
        i = sectors.index(sec)
        a, b = int(M.off[i]), int(M.off[i + 1])
        print("  sector (kx, ky)=%s (%d states): basis, oracle rows" % (sec, b - a), flush=True)
        ref = O.MomentumFermions(n, nx, ny, sec[0], sec[1])
        assert ref.dim == b - a
        assert np.array_equal(states[a:b], ref.states())
        xr = download(arch, dx, 2 * a * k, 2 * (b - a) * k)
        X = xr[0::2] + 1j * xr[1::2]
        for r0, r1 in slices(b - a, pieces=2):
            yr = download(arch, dy, 2 * (a + r0) * k, 2 * (r1 - r0) * k)
            Y = yr[0::2] + 1j * yr[1::2]
            Yr = O.fm_hv_rows(ref, fac, X, r0, r1, k=k, threads=THREADS)
            for v in range(k):
                yv, rv = Y[v::k], Yr[v::k]
                assert np.linalg.norm(yv - rv) <= 1e-12 * np.linalg.norm(rv), (sec, r0, r1, v)
    dx.free()
    dy.free()
    M.close()
