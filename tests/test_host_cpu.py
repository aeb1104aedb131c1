"""CPU checks of the product's host layer and of the C ABI boundary (no GPU calls)."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(dhh?_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.mark.parametrize("header,lib", [("dhgpu.h", "libdhgpu.so"), ("diagham_host.h", "libdiagham_host.so")])
def test_library_exports_every_declared_symbol(header, lib):
    from diagham_amd import capi
    path = os.path.join(ROOT, "diagham_amd", lib)
    assert os.path.exists(path), "build the native libraries first"
    out = subprocess.run(["nm", "-D", "--defined-only", path], stdout=subprocess.PIPE, text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if " T " in line)
    decl = declared_functions(header)
    assert decl, header
    missing = [d for d in decl if d not in exported]
    assert not missing, missing
    # the Python bindings cover exactly the declared ABI
    sigs = capi.GPU_SIGNATURES if lib == "libdhgpu.so" else capi.HOST_SIGNATURES
    assert sorted(sigs) == decl


def test_libraries_load_without_gpu():
    from diagham_amd import capi
    capi.gpu()
    capi.host()


def test_null_handles_are_refused_without_gpu():
    """Entry points given a null handle or buffer return DH_E_ARG before touching a device."""
    from diagham_amd import capi
    g = capi.gpu()
    E = -1  # DH_E_ARG
    assert g.dh_sector_set_launch(None, 1, 256) == E
    assert g.dh_sector_set_blas_launch(None, 8, 0) == E
    assert g.dh_sectors_hv(None, None, None, 1, 0, None) == E
    assert g.dh_sectors_block_dots(None, None, 4, None, 1, None, None) == E
    assert g.dh_sectors_block_axpy(None, None, 4, None, 1, None, None) == E
    assert g.dh_sectors_block_rmul(None, None, 1, None, None) == E
    assert g.dh_sector_set_destroy(None) == E
    assert g.dh_hv(None, None, None, 1, 0, 0, -1) == E
    assert g.dh_block_dots(None, None, 4, None, 16, 1, None) == E
    assert g.dh_block_axpy(None, None, 4, None, 16, 1, None) == E


def test_dimension_without_gpu(golden):
    from diagham_amd import capi
    assert capi.fermion_sphere_dimension(8, 0, 21) == golden["fermions_sphere_n8_2s21_lz0"]["dimension"]
    assert capi.fermion_sphere_dimension(10, 0, 27) == golden["fermions_sphere_n10_2s27_lz0"]["dimension"]
    assert capi.fermion_sphere_dimension(12, 0, 33) == 7764392
    assert capi.fermion_sphere_dimension(8, 1, 21) == 0  # wrong parity


def test_momentum_sector_dimension_without_gpu():
    """dh_momentum_sector_dimension against the oracle's basis of FermionOnSquareLatticeMomentumSpace
    (4 fermions on 4 x 3 sites, every sector), the sum over the Brillouin zone, the sizes of the
    30 sectors of N=10 on 6 x 5 sites (1,001,400 to 1,001,603 states) and its argument errors."""
    from math import comb
    import oracle_lib as O
    from diagham_amd import capi
    dim = capi.gpu().dh_momentum_sector_dimension
    total = 0
    for kx in range(4):
        for ky in range(3):
            d = dim(4, 4, 3, kx, ky)
            assert d == O.MomentumFermions(4, 4, 3, kx, ky).dim, (kx, ky)
            total += d
    assert total == comb(12, 4)
    big = [dim(10, 6, 5, kx, ky) for kx in range(6) for ky in range(5)]
    assert sum(big) == comb(30, 10) == 30045015 and min(big) == 1001400 and max(big) == 1001603
    assert dim(4, 4, 3, 4, 0) == -1 and dim(0, 4, 3, 0, 0) == -1 and dim(13, 4, 3, 0, 0) == -1


@pytest.mark.parametrize("lz", [8, 21, 27, 33])
def test_host_pseudopotentials_match_oracle(lz):
    from diagham_amd import capi
    a = capi.coulomb_pseudopotentials(lz, round_to_file=False)
    b = O.coulomb_pseudopotentials(lz)
    assert np.max(np.abs(a - b) / np.abs(b)) < 1e-13
    assert np.array_equal(capi.coulomb_pseudopotentials(lz, True), O.round14(a))


@pytest.mark.parametrize("lz", [8, 21, 27, 33])
def test_host_factors_match_oracle(lz):
    """Product factors (113-bit Racah CG) vs oracle factors (reference recursion)."""
    from diagham_amd import capi
    V = O.round14(O.coulomb_pseudopotentials(lz))
    f, off = capi.sphere_factors(lz, V)
    m1, m2, nm3, m3, F = O.Factors(lz, V).tables()
    assert f.size == F.size
    assert np.array_equal(np.diff(off), nm3)
    assert np.max(np.abs(f - F)) < 1e-13 * np.max(np.abs(F))


def test_channel_split_reconstructs_factors():
    """sum_c V_c (-4) cg[a][c] cg[b][c] == factor(a -> b): the adjoint split is exact."""
    from diagham_amd import capi
    lz = 15
    rng = np.random.default_rng(1)
    V = rng.uniform(0.1, 1.0, lz + 1)
    f, off = capi.sphere_factors(lz, V)
    mom, tab = capi.sphere_channel_table(lz)
    assert list(mom) == list(range(1, lz + 1, 2))
    w = V[mom]
    pos = 0
    for m1 in range(lz + 1):
        for m2 in range(m1):
            a = m1 * (m1 - 1) // 2 + m2
            s = m1 + m2
            for m3 in range(s // 2 + 1, min(s, lz) + 1):
                m4 = s - m3
                b = m3 * (m3 - 1) // 2 + m4
                rec = -4.0 * np.sum(w * tab[a] * tab[b])
                assert abs(rec - f[pos]) < 1e-13
                pos += 1
    assert pos == f.size


def test_host_clebsch_gordan_matches_oracle():
    from diagham_amd import capi
    cg = O.lib().or_cg_create(27, 27)
    rng = np.random.default_rng(3)
    for _ in range(500):
        m1 = int(rng.integers(0, 28)) * 2 - 27
        m2 = int(rng.integers(0, 28)) * 2 - 27
        jmin = abs(m1 + m2)
        J = jmin + 2 * int(rng.integers(0, (54 - jmin) // 2 + 1))
        assert abs(capi.clebsch_gordan(27, 27, m1, m2, J) - O.lib().or_cg_get(cg, m1, m2, J)) < 1e-13
    O.lib().or_cg_destroy(cg)


@pytest.mark.parametrize("n,b", [(50, 1), (64, 4), (96, 8), (33, 16)])
def test_band_eigenvalues(n, b):
    from diagham_amd import capi
    rng = np.random.default_rng(n + b)
    A = rng.standard_normal((n, n))
    A = A + A.T
    A[np.abs(np.subtract.outer(np.arange(n), np.arange(n))) > b] = 0.0
    assert np.max(np.abs(capi.band_eigenvalues(A, b) - np.linalg.eigvalsh(A))) < 1e-12 * np.max(np.abs(A)) * n


def test_pseudopotential_file_reader(tmp_path):
    from diagham_amd import capi
    import ctypes
    lz = 9
    v = capi.coulomb_pseudopotentials(lz)
    p = tmp_path / "pp.dat"
    p.write_text("# pseudopotentials on the sphere\nName = test\nPseudopotentials = %s\n" %
                 " ".join("%.14g" % x for x in v))
    out = np.zeros(lz + 1)
    rc = capi.host().dhh_read_pseudopotentials(str(p).encode(), lz, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0 and np.array_equal(out, v)


def test_host_checkerboard_band_and_factors_match_oracle():
    """Product host tables of the Chern-band model vs the oracle restatement."""
    import ctypes
    from diagham_amd import capi
    nx, ny = 6, 5
    u = np.zeros(nx * ny * 4)
    e = np.zeros(nx * ny * 2)
    assert capi.host().dhh_checkerboard_lower_band(nx, ny, 1.0, 0.5, -0.5, 0.0, 0.0, 0.0, capi.dptr(u), capi.dptr(e)) == 0
    uo, eo = O.checkerboard_lower_band(nx, ny, 1.0, 0.5, -0.5)
    assert np.max(np.abs((u[0::2] + 1j * u[1::2]).reshape(-1, 2) - uo)) < 1e-14
    assert np.max(np.abs(e.reshape(-1, 2) - eo)) < 1e-14
    npairs = nx * ny * (nx * ny - 1) // 2
    cnt = np.zeros(nx * ny, np.int32)
    pairs = np.zeros(2 * npairs, np.int32)
    f = np.zeros(2 * npairs * npairs)
    npr, nf = ctypes.c_int(npairs), ctypes.c_int64(npairs * npairs)
    assert capi.host().dhh_checkerboard_factors(nx, ny, 1.0, 0.5, -0.5, 0.0, 1.0, 0.0, 1, cnt.ctypes.data_as(capi.P_I),
                                                pairs.ctypes.data_as(capi.P_I), ctypes.byref(npr), capi.dptr(f),
                                                ctypes.byref(nf)) == 0
    fac = O.LatticeFactors(nx, ny, uo)
    pos, fpos = 0, 0
    for s in range(nx * ny):
        p_ref, F_ref = fac.sector(s)
        m = cnt[s]
        assert m == len(p_ref)
        assert np.array_equal(pairs[2 * pos:2 * (pos + m)].reshape(m, 2), p_ref)
        F = (f[2 * fpos:2 * (fpos + m * m):2] + 1j * f[2 * fpos + 1:2 * (fpos + m * m):2]).reshape(m, m)
        assert np.max(np.abs(F - F_ref)) < 1e-14
        assert np.max(np.abs(F - F.conj().T)) < 1e-14  # Hermitian factor matrix per momentum sum
        pos += m
        fpos += m * m


def test_sector_solve_settings_are_checked():
    """dhh_sectors_launch and dhh_sectors_alone_rows refuse shapes and thresholds outside their
    ranges with a message, and accept the reset values (no device needed)."""
    from diagham_amd import capi
    h = capi.host()
    assert h.dhh_sectors_launch(3, 256, capi.INDEX_RANK) == capi.DH_E_ARG
    assert b"lanes per row" in h.dhh_last_error()
    assert h.dhh_sectors_launch(2, 100, capi.INDEX_RANK) == capi.DH_E_ARG
    assert h.dhh_sectors_launch(0, 0, 0) == 0
    assert h.dhh_sectors_alone_rows(-5) == capi.DH_E_ARG
    assert b"alone rows" in h.dhh_last_error()
    assert h.dhh_sectors_alone_rows(0) == 0
    assert h.dhh_sectors_alone_rows(2000000) == 0
