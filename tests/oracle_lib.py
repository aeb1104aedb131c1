"""ctypes view of the CPU oracle (oracle/liboracle.so) -- test infrastructure only.

Only tests, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg use this
module; the product path never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB_PATH = os.path.join(_ROOT, "oracle", "liboracle.so")
_lib = None

c_int, c_long, c_double, c_void_p, c_uint64, c_ulong = (
    ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_ulong)
P_D = ctypes.POINTER(ctypes.c_double)
P_L = ctypes.POINTER(ctypes.c_long)
P_I = ctypes.POINTER(ctypes.c_int)
P_U64 = ctypes.POINTER(ctypes.c_uint64)


def _sig(lib, name, restype, argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = argtypes


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", os.path.dirname(_LIB_PATH)])
    L = ctypes.CDLL(_LIB_PATH)
    _sig(L, "or_fs_dimension", c_long, [c_int, c_int, c_int])
    _sig(L, "or_fs_create", c_void_p, [c_int, c_int, c_int, c_ulong])
    _sig(L, "or_fs_destroy", None, [c_void_p])
    _sig(L, "or_fs_dim", c_long, [c_void_p])
    _sig(L, "or_fs_states", P_U64, [c_void_p])
    _sig(L, "or_fs_state_lzmax", P_I, [c_void_p])
    _sig(L, "or_fs_find_state_index", c_long, [c_void_p, c_uint64, c_int])
    _sig(L, "or_fs_adadaa", c_long, [c_void_p, c_long, c_int, c_int, c_int, c_int, P_D])
    _sig(L, "or_fs_aa", c_double, [c_void_p, c_long, c_int, c_int])
    _sig(L, "or_fs_adad", c_long, [c_void_p, c_int, c_int, P_D])
    _sig(L, "or_fs_prod_a", c_double, [c_void_p, c_long, P_I, c_int])
    _sig(L, "or_fs_prod_ad", c_long, [c_void_p, P_I, c_int, P_D])
    _sig(L, "or_nbody_hv", None, [c_void_p, c_int, c_int, P_I, P_I, P_D, P_D, P_D])
    _sig(L, "or_cg_create", c_void_p, [c_int, c_int])
    _sig(L, "or_cg_destroy", None, [c_void_p])
    _sig(L, "or_cg_get", c_double, [c_void_p, c_int, c_int, c_int])
    _sig(L, "or_coulomb_pseudopotentials", None, [c_int, P_D])
    _sig(L, "or_coulomb_closed_form", None, [c_int, P_D])
    _sig(L, "or_factors_create", c_void_p, [c_int, P_D])
    _sig(L, "or_factors_hermitian", None, [c_void_p])
    _sig(L, "or_factors_destroy", None, [c_void_p])
    _sig(L, "or_factors_nbr_m12", c_int, [c_void_p])
    _sig(L, "or_factors_nbr", c_int, [c_void_p])
    _sig(L, "or_factors_get", None, [c_void_p, P_I, P_I, P_I, P_I, P_D])
    _sig(L, "or_fs_hv", None, [c_void_p, c_void_p, P_D, P_D, c_double, c_int])
    _sig(L, "or_fs_hv_smp", None, [c_void_p, c_void_p, P_D, P_D, c_double, c_int, c_int])
    _sig(L, "or_fs_row_elements", c_long, [c_void_p, c_void_p, c_long, P_L, P_D, c_long])
    _sig(L, "or_fs_count_elements", c_long, [c_void_p, c_void_p])
    _sig(L, "or_fs_hv_rows", None, [c_void_p, c_void_p, P_D, P_D, c_double, c_long, c_long, c_int, c_int])
    _sig(L, "or_fs_csr_create", c_void_p, [c_void_p, c_void_p, c_int])
    _sig(L, "or_csr_destroy", None, [c_void_p])
    _sig(L, "or_csr_nnz", c_long, [c_void_p])
    _sig(L, "or_csr_hv", None, [c_void_p, P_D, P_D, c_int])
    _sig(L, "or_spin_create", c_void_p, [c_int, c_int, c_int])
    _sig(L, "or_spin_destroy", None, [c_void_p])
    _sig(L, "or_spin_dim", c_long, [c_void_p])
    _sig(L, "or_spin_states", P_U64, [c_void_p])
    _sig(L, "or_spin_find_state_index", c_long, [c_void_p, c_uint64])
    _sig(L, "or_spin_hv", None, [c_void_p, P_D, P_D, c_int, P_D, P_D])
    _sig(L, "or_spin_hv_smp", None, [c_void_p, P_D, P_D, c_int, P_D, P_D, c_int])
    _sig(L, "or_spin_hv_rows", None, [c_void_p, P_D, P_D, c_int, P_D, P_D, c_long, c_long, c_int])
    _sig(L, "or_tridiag_eigenvalues", c_int, [c_int, P_D, P_D, c_int])
    _sig(L, "or_fs_lanczos", c_int, [c_void_p, c_void_p, c_double, c_int, c_int, c_int, c_double,
                                     c_uint64, c_int, P_D, P_D])
    _sig(L, "or_spin_lanczos", c_int, [c_void_p, P_D, P_D, c_int, c_int, c_int, c_double, c_uint64,
                                       c_int, P_D])
    _sig(L, "or_splitmix_fill", None, [c_uint64, P_D, c_long])
    _lib = L
    return L


def dptr(a):
    return a.ctypes.data_as(P_D)


def splitmix(seed, n):
    out = np.empty(n, dtype=np.float64)
    lib().or_splitmix_fill(seed, dptr(out), n)
    return out


class FermionSphere:
    """FermionOnSphere restatement (FQHE/src/HilbertSpace/FermionOnSphere.cc)."""

    def __init__(self, n, total_lz, lz_max, memory=9 << 20):
        self.n, self.total_lz, self.lz_max = n, total_lz, lz_max
        self.h = lib().or_fs_create(n, total_lz, lz_max, memory)
        if not self.h:
            raise ValueError("empty Hilbert space")
        self.dim = lib().or_fs_dim(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_fs_destroy(self.h)
            self.h = None

    def states(self):
        p = lib().or_fs_states(self.h)
        return np.ctypeslib.as_array(p, shape=(self.dim,)).copy()

    def state_lzmax(self):
        p = lib().or_fs_state_lzmax(self.h)
        return np.ctypeslib.as_array(p, shape=(self.dim,)).copy()

    def find(self, state, lzmax):
        return lib().or_fs_find_state_index(self.h, int(state), int(lzmax))

    def adadaa(self, i, m1, m2, n1, n2):
        c = c_double(0.0)
        idx = lib().or_fs_adadaa(self.h, i, m1, m2, n1, n2, ctypes.byref(c))
        return idx, c.value

    def aa(self, i, n1, n2):
        return lib().or_fs_aa(self.h, i, n1, n2)

    def adad(self, m1, m2):
        c = c_double(0.0)
        idx = lib().or_fs_adad(self.h, m1, m2, ctypes.byref(c))
        return idx, c.value

    def prod_a(self, i, n):
        n = np.ascontiguousarray(n, dtype=np.int32)
        return lib().or_fs_prod_a(self.h, i, n.ctypes.data_as(P_I), n.size)

    def prod_ad(self, m):
        m = np.ascontiguousarray(m, dtype=np.int32)
        c = c_double(0.0)
        idx = lib().or_fs_prod_ad(self.h, m.ctypes.data_as(P_I), m.size, ctypes.byref(c))
        return idx, c.value


class Factors:
    """Interaction factor tables of ParticleOnSphereGenericHamiltonian (fermions)."""

    def __init__(self, lz_max, pseudopotentials, hermitian=False):
        v = np.ascontiguousarray(pseudopotentials, dtype=np.float64)
        assert v.size == lz_max + 1
        self.h = lib().or_factors_create(lz_max, dptr(v))
        if hermitian:
            lib().or_factors_hermitian(self.h)
        self.lz_max = lz_max

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_factors_destroy(self.h)
            self.h = None

    def tables(self):
        n12 = lib().or_factors_nbr_m12(self.h)
        nf = lib().or_factors_nbr(self.h)
        m1 = np.zeros(n12, np.int32)
        m2 = np.zeros(n12, np.int32)
        nm3 = np.zeros(n12, np.int32)
        m3 = np.zeros(max(nf, 1), np.int32)
        f = np.zeros(max(nf, 1), np.float64)
        lib().or_factors_get(self.h, m1.ctypes.data_as(P_I), m2.ctypes.data_as(P_I),
                             nm3.ctypes.data_as(P_I), m3.ctypes.data_as(P_I), dptr(f))
        return m1, m2, nm3, m3[:nf], f[:nf]

    @property
    def nbr_factors(self):
        return lib().or_factors_nbr(self.h)


def coulomb_pseudopotentials(nbr_flux):
    out = np.zeros(nbr_flux + 1)
    lib().or_coulomb_pseudopotentials(nbr_flux, dptr(out))
    return out


def coulomb_closed_form(nbr_flux):
    out = np.zeros(nbr_flux + 1)
    lib().or_coulomb_closed_form(nbr_flux, dptr(out))
    return out


def round14(v):
    """Values as written by CoulombPseudopotentials (File.precision(14), %.14g)."""
    return np.array([float("%.14g" % x) for x in v])


def fs_hv(space, factors, x, shift=0.0, hermitian=False, threads=1):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    if threads > 1:
        lib().or_fs_hv_smp(space.h, factors.h, dptr(x), dptr(y), shift, int(hermitian), threads)
    else:
        lib().or_fs_hv(space.h, factors.h, dptr(x), dptr(y), shift, int(hermitian))
    return y


def fs_hv_rows(space, factors, x, first, last, shift=0.0, k=1, threads=8):
    """rows [first, last) of (H + shift) X for k vectors in [row][k] layout (x of dim * k doubles),
    gather form (or_fs_hv_rows); returns (last - first) * k doubles"""
    x = np.ascontiguousarray(x, dtype=np.float64)
    assert x.size == space.dim * k
    out = np.zeros((last - first) * k)
    lib().or_fs_hv_rows(space.h, factors.h, dptr(x), dptr(out), shift, first, last, k, threads)
    return out


class StoredRows:
    """The rows of H (or_fs_csr_create) for repeated products (or_csr_hv)."""

    def __init__(self, space, factors, threads=8):
        self.h = lib().or_fs_csr_create(space.h, factors.h, threads)
        self.dim, self.threads = space.dim, threads
        self.nnz = lib().or_csr_nnz(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_csr_destroy(self.h)
            self.h = None

    def hv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        lib().or_csr_hv(self.h, dptr(x), dptr(y), self.threads)
        return y


def nbody_tuples(nbr_orbitals, nbody):
    """Orbital tuples (descending within a tuple) grouped by orbital sum, in the order of
    AbstractQHEOnSphereNBodyInteractionHamiltonian::GetAllSkewSymmetricIndices (:1600):
    lexicographically descending.  Returns (per-sum counts, flat tuple array)."""
    import itertools
    combos = sorted(itertools.combinations(range(nbr_orbitals), nbody), reverse=True)
    combos = [tuple(sorted(c, reverse=True)) for c in combos]
    combos.sort(reverse=True)
    nsums = nbody * (nbr_orbitals - 1) + 1
    per = [[] for _ in range(nsums)]
    for c in combos:
        per[sum(c)].append(c)
    counts = np.array([len(p) for p in per], np.int32)
    flat = np.array([x for p in per for c in p for x in c], np.int32)
    return counts, flat


def random_nbody_factors(counts, seed):
    """Random symmetric factor matrix per orbital sum, concatenated."""
    rng = np.random.default_rng(seed)
    out = []
    for c in counts:
        a = rng.standard_normal((c, c))
        out.append(((a + a.T) / 2).ravel())
    return np.concatenate(out) if out else np.zeros(0)


def nbody_hv(space, nbody, counts, tuples, factors, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    counts = np.ascontiguousarray(counts, dtype=np.int32)
    tuples = np.ascontiguousarray(tuples, dtype=np.int32)
    factors = np.ascontiguousarray(factors, dtype=np.float64)
    lib().or_nbody_hv(space.h, nbody, counts.size, counts.ctypes.data_as(P_I), tuples.ctypes.data_as(P_I),
                      dptr(factors), dptr(x), dptr(y))
    return y


def fs_row_elements(space, factors, row, max_out=100000):
    t = np.zeros(max_out, np.int64)
    c = np.zeros(max_out, np.float64)
    n = lib().or_fs_row_elements(space.h, factors.h, row, t.ctypes.data_as(P_L), dptr(c), max_out)
    assert n <= max_out
    return t[:n], c[:n]


def fs_count_elements(space, factors):
    return lib().or_fs_count_elements(space.h, factors.h)


def fs_lanczos(space, factors, n_eig, shift=0.0, hermitian=False, max_iter=3000, precision=1e-14,
               seed=0x5EED0000, threads=8):
    ev = np.zeros(n_eig)
    it = lib().or_fs_lanczos(space.h, factors.h, shift, int(hermitian), n_eig, max_iter, precision,
                             seed, threads, dptr(ev), None)
    return ev, it


class SpinChain:
    """Spin1_2Chain restatement (Spin/src/HilbertSpace/Spin1_2Chain.cc)."""

    def __init__(self, length, sz, memory=1000000):
        self.length = length
        self.h = lib().or_spin_create(length, sz, memory)
        self.dim = lib().or_spin_dim(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_spin_destroy(self.h)
            self.h = None

    def states(self):
        p = lib().or_spin_states(self.h)
        return np.ctypeslib.as_array(p, shape=(self.dim,)).copy()

    def find(self, state):
        return lib().or_spin_find_state_index(self.h, int(state))

    def couplings(self, j=1.0, jz=1.0):
        return np.full(self.length, j, np.float64), np.full(self.length, jz, np.float64)

    def hv(self, x, j=1.0, jz=1.0, periodic=True, threads=1):
        J, Jz = self.couplings(j, jz)
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        if threads > 1:
            lib().or_spin_hv_smp(self.h, dptr(J), dptr(Jz), int(periodic), dptr(x), dptr(y), threads)
        else:
            lib().or_spin_hv(self.h, dptr(J), dptr(Jz), int(periodic), dptr(x), dptr(y))
        return y

    def lanczos(self, n_eig, j=1.0, jz=1.0, periodic=True, max_iter=3000, precision=1e-14,
                seed=0x5EED0000, threads=8):
        J, Jz = self.couplings(j, jz)
        ev = np.zeros(n_eig)
        it = lib().or_spin_lanczos(self.h, dptr(J), dptr(Jz), int(periodic), n_eig, max_iter,
                                   precision, seed, threads, dptr(ev))
        return ev, it


def _bind_momentum(L):
    if getattr(L, "_momentum_bound", False):
        return L
    _sig(L, "or_fm_create", c_void_p, [c_int, c_int, c_int, c_int, c_int, c_ulong])
    _sig(L, "or_checkerboard_lower_band", None, [c_int, c_int, c_double, c_double, c_double, c_double, c_double,
                                                 c_double, P_D, P_D])
    _sig(L, "or_fm_factors_create", c_void_p, [c_int, c_int, P_D, c_double, c_double, c_int])
    _sig(L, "or_fm_factors_destroy", None, [c_void_p])
    _sig(L, "or_fm_factors_sector_size", c_int, [c_void_p, c_int])
    _sig(L, "or_fm_factors_get", None, [c_void_p, c_int, P_I, P_D, P_D])
    _sig(L, "or_fm_hv", None, [c_void_p, c_void_p, P_D, P_D, c_int])
    _sig(L, "or_fm_row_elements", c_long, [c_void_p, c_void_p, c_long, P_L, P_D, P_D, c_long])
    _sig(L, "or_fm_hv_rows", None, [c_void_p, c_void_p, P_D, P_D, c_long, c_long, c_int, c_int])
    _sig(L, "or_torus_factors_create", c_void_p, [c_int, c_double, c_int, P_D, c_int])
    _sig(L, "or_ftt_create", c_void_p, [c_int, c_int, c_int, c_int])
    _sig(L, "or_ftt_destroy", None, [c_void_p])
    _sig(L, "or_ftt_dim", c_long, [c_void_p])
    _sig(L, "or_ftt_states", P_U64, [c_void_p])
    _sig(L, "or_ftt_orbits", P_I, [c_void_p])
    _sig(L, "or_ftt_reorder", P_U64, [c_void_p])
    _sig(L, "or_ftt_modulo", c_int, [c_void_p])
    _sig(L, "or_ftt_hv", None, [c_void_p, c_void_p, P_D, P_D])
    L._momentum_bound = True
    return L


def checkerboard_lower_band(nx, ny, t1=1.0, t2=0.5, tpp=-0.5, mus=0.0, gx=0.0, gy=0.0):
    """OneBodyBasis[k][0][alpha] (lower band) as complex (nx*ny, 2), band energies (nx*ny, 2)."""
    L = _bind_momentum(lib())
    u = np.zeros(nx * ny * 4)
    e = np.zeros(nx * ny * 2)
    L.or_checkerboard_lower_band(nx, ny, t1, t2, tpp, mus, gx, gy, dptr(u), dptr(e))
    return (u[0::2] + 1j * u[1::2]).reshape(nx * ny, 2), e.reshape(nx * ny, 2)


class MomentumFermions:
    """FermionOnSquareLatticeMomentumSpace restatement."""

    def __init__(self, n, nx, ny, kx, ky, memory=9 << 20):
        L = _bind_momentum(lib())
        self.n, self.nx, self.ny, self.kx, self.ky = n, nx, ny, kx, ky
        self.h = L.or_fm_create(n, nx, ny, kx, ky, memory)
        if not self.h:
            raise ValueError("empty Hilbert space")
        self.dim = L.or_fs_dim(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_fs_destroy(self.h)
            self.h = None

    def states(self):
        p = lib().or_fs_states(self.h)
        return np.ctypeslib.as_array(p, shape=(self.dim,)).copy()


class LatticeFactors:
    """ParticleOnLatticeCheckerboardLatticeSingleBandHamiltonian factor tables (fermions)."""

    def __init__(self, nx, ny, u, u_potential=1.0, v_potential=0.0, flat_band=True):
        L = _bind_momentum(lib())
        uu = np.zeros(nx * ny * 4)
        uu[0::2] = np.asarray(u).reshape(-1).real
        uu[1::2] = np.asarray(u).reshape(-1).imag
        self.nx, self.ny = nx, ny
        self.h = L.or_fm_factors_create(nx, ny, dptr(uu), u_potential, v_potential, int(flat_band))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_fm_factors_destroy(self.h)
            self.h = None

    def sector(self, s):
        L = lib()
        n = L.or_fm_factors_sector_size(self.h, s)
        pairs = np.zeros(2 * n, np.int32)
        fr = np.zeros(n * n + 1)
        fi = np.zeros(n * n + 1)
        L.or_fm_factors_get(self.h, s, pairs.ctypes.data_as(P_I), dptr(fr), dptr(fi))
        return pairs.reshape(n, 2), (fr[:n * n] + 1j * fi[:n * n]).reshape(n, n)


def fm_hv(space, factors, x, hermitian=False):
    """H x for a complex vector x (numpy complex128)."""
    xr = np.empty(2 * space.dim)
    xr[0::2], xr[1::2] = np.real(x), np.imag(x)
    y = np.zeros(2 * space.dim)
    lib().or_fm_hv(space.h, factors.h, dptr(xr), dptr(y), int(hermitian))
    return y[0::2] + 1j * y[1::2]


def fm_hv_rows(space, factors, x, first, last, k=1, threads=8):
    """rows [first, last) of H X for k complex vectors (x: complex128 of dim * k, [row][k]),
    gather form (or_fm_hv_rows); returns (last - first) * k complex values"""
    L = _bind_momentum(lib())
    x = np.asarray(x)
    assert x.size == space.dim * k
    xr = np.empty(2 * x.size)
    xr[0::2], xr[1::2] = np.real(x).ravel(), np.imag(x).ravel()
    out = np.zeros(2 * (last - first) * k)
    L.or_fm_hv_rows(space.h, factors.h, dptr(xr), dptr(out), first, last, k, threads)
    return out[0::2] + 1j * out[1::2]


def fm_row_elements(space, factors, row, max_out=20000):
    t = np.zeros(max_out, np.int64)
    cr = np.zeros(max_out)
    ci = np.zeros(max_out)
    n = lib().or_fm_row_elements(space.h, factors.h, row, t.ctypes.data_as(P_L), dptr(cr), dptr(ci), max_out)
    assert n <= max_out
    return t[:n], cr[:n] + 1j * ci[:n]


class TorusFactors(LatticeFactors):
    """ParticleOnTorusGenericHamiltonian (pseudopotentials) or ParticleOnTorusCoulombHamiltonian
    (coulomb_level >= 0) factor tables, in the momentum-space layout with Nx = 1."""

    def __init__(self, nflux, ratio=1.0, pseudo=(), coulomb_level=-1):
        L = _bind_momentum(lib())
        pp = np.ascontiguousarray(pseudo, dtype=np.float64)
        if pp.size == 0:
            pp = np.zeros(1)
        self.nx, self.ny = 1, nflux
        self.h = L.or_torus_factors_create(nflux, ratio, int(len(pseudo)), dptr(pp), coulomb_level)


def torus_space(n, nflux, ky):
    """FermionOnTorus(n, nflux, ky): the momentum-space basis with Nx = 1."""
    return MomentumFermions(n, 1, nflux, 0, ky)


def torus_states_brute_force(n, nflux, ky):
    """Every n-subset of nflux orbitals with total momentum ky mod nflux, descending."""
    from itertools import combinations
    out = [sum(1 << o for o in c) for c in combinations(range(nflux), n) if sum(c) % nflux == ky % nflux]
    return np.array(sorted(out, reverse=True), dtype=np.uint64)


def dense_fm(space, factors):
    """Dense H of a momentum-space (or torus) sector from the oracle product (small spaces)."""
    d = space.dim
    H = np.zeros((d, d), complex)
    for c in range(d):
        e = np.zeros(d, complex)
        e[c] = 1.0
        H[:, c] = fm_hv(space, factors, e)
    return H


class TorusTranslations:
    """FermionOnTorusWithMagneticTranslations(n, nflux, kx, ky) restatement."""

    def __init__(self, n, nflux, kx, ky):
        L = _bind_momentum(lib())
        self.n, self.nflux, self.kx, self.ky = n, nflux, kx, ky
        self.h = L.or_ftt_create(n, nflux, kx, ky)
        if not self.h:
            raise ValueError("bad torus space")
        self.dim = L.or_ftt_dim(self.h)
        self.modulo = L.or_ftt_modulo(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_ftt_destroy(self.h)
            self.h = None

    def states(self):
        return np.ctypeslib.as_array(lib().or_ftt_states(self.h), shape=(max(self.dim, 1),))[:self.dim].copy()

    def orbits(self):
        return np.ctypeslib.as_array(lib().or_ftt_orbits(self.h), shape=(max(self.dim, 1),))[:self.dim].copy()

    def reorder(self):
        return np.ctypeslib.as_array(lib().or_ftt_reorder(self.h), shape=(max(self.dim, 1),))[:self.dim].copy()

    def hv(self, factors, x):
        xr = np.empty(2 * self.dim)
        xr[0::2], xr[1::2] = np.real(x), np.imag(x)
        y = np.zeros(2 * self.dim)
        lib().or_ftt_hv(self.h, factors.h, dptr(xr), dptr(y))
        return y[0::2] + 1j * y[1::2]

    def dense(self, factors):
        H = np.zeros((self.dim, self.dim), complex)
        for c in range(self.dim):
            e = np.zeros(self.dim, complex)
            e[c] = 1.0
            H[:, c] = self.hv(factors, e)
        return H
