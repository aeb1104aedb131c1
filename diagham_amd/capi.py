"""ctypes bindings of the native libraries (include/dhgpu.h, include/diagham_host.h).

The product path is native: importing this module loads ``libdhgpu.so`` (HIP kernels for
gfx950) and ``libdiagham_host.so`` (C++ host layer) from this directory and raises if they
are missing.  There is no Python or CPU fallback.
"""
import atexit
import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_GPU = os.path.join(_HERE, "libdhgpu.so")
LIB_HOST = os.path.join(_HERE, "libdiagham_host.so")

DH_OK, DH_E_ARG, DH_E_HIP, DH_E_NOMEM, DH_E_STATE, DH_E_RANGE = 0, -1, -2, -3, -4, -5
INDEX_BSEARCH, INDEX_RANK, INDEX_RANK_X4, INDEX_RANK_FREE, INDEX_RANK_RANGE = 0, 1, 2, 3, 4

c_int, c_int64, c_double, c_void_p, c_uint64, c_size_t, c_char_p = (
    ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_size_t,
    ctypes.c_char_p)
P_D = ctypes.POINTER(ctypes.c_double)
P_I = ctypes.POINTER(ctypes.c_int)
P_I64 = ctypes.POINTER(ctypes.c_int64)
P_U64 = ctypes.POINTER(ctypes.c_uint64)
PP = ctypes.POINTER(ctypes.c_void_p)

_gpu = None
_host = None

# name: (restype, argtypes) of every entry point of include/dhgpu.h
GPU_SIGNATURES = {
    "dh_create": (c_int, [PP, c_int]),
    "dh_destroy": (c_int, [c_void_p]),
    "dh_last_error": (c_char_p, [c_void_p]),
    "dh_set_stream": (c_int, [c_void_p, c_void_p]),
    "dh_synchronize": (c_int, [c_void_p]),
    "dh_device_info": (c_int, [c_void_p, P_I, c_char_p]),
    "dh_memory_info": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "dh_memory_peak": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_size_t), c_int]),
    "dh_malloc": (c_int, [c_void_p, PP, c_size_t]),
    "dh_free": (c_int, [c_void_p, c_void_p]),
    "dh_memcpy_h2d": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "dh_memcpy_d2h": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "dh_memset": (c_int, [c_void_p, c_void_p, c_int, c_size_t]),
    "dh_fermion_sphere_create": (c_int, [c_void_p, PP, c_int, c_int, c_int]),
    "dh_fermion_sphere_dimension": (c_int64, [c_int, c_int, c_int]),
    "dh_fermion_sphere_from_states": (c_int, [c_void_p, PP, c_int, c_int, c_int, P_U64, c_int64]),
    "dh_spin_chain_create": (c_int, [c_void_p, PP, c_int, c_int]),
    "dh_basis_destroy": (c_int, [c_void_p]),
    "dh_basis_dimension": (c_int64, [c_void_p]),
    "dh_basis_get_states": (c_int, [c_void_p, P_U64]),
    "dh_basis_find_state_index": (c_int, [c_void_p, P_U64, c_int64, c_int, P_I64]),
    "dh_sphere_hamiltonian_create": (c_int, [c_void_p, PP, c_void_p, P_D, P_I, c_int]),
    "dh_sphere_nbody_hamiltonian_create": (c_int, [c_void_p, PP, c_void_p, c_int, P_I, c_int, P_I, P_D]),
    "dh_sphere_hamiltonian_set_channels": (c_int, [c_void_p, P_D, c_int, P_D]),
    "dh_spin_chain_hamiltonian_create": (c_int, [c_void_p, PP, c_void_p, P_D, P_D, c_int]),
    "dh_hamiltonian_destroy": (c_int, [c_void_p]),
    "dh_hamiltonian_set_shift": (c_int, [c_void_p, c_double]),
    "dh_hamiltonian_dimension": (c_int64, [c_void_p]),
    "dh_hamiltonian_set_launch": (c_int, [c_void_p, c_int, c_int, c_int, c_int]),
    "dh_hv": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int64, c_int64]),
    "dh_hv_vectors": (c_int, [c_void_p, PP, PP, c_int, c_int, c_int64, c_int64]),
    "dh_hamiltonian_count_elements": (c_int, [c_void_p, P_I64]),
    "dh_hamiltonian_precompute": (c_int, [c_void_p, c_int64, P_I64]),
    "dh_hamiltonian_release_precomputed": (c_int, [c_void_p]),
    "dh_hamiltonian_row_elements": (c_int, [c_void_p, c_int64, P_I64, P_D, c_int64, P_I64]),
    "dh_sphere_adjoint": (c_int, [c_void_p, c_void_p, c_void_p, P_D, c_void_p]),
    "dh_sector_set_create": (c_int, [c_void_p, PP, c_int, c_int, P_I, c_int]),
    "dh_sector_set_destroy": (c_int, [c_void_p]),
    "dh_sector_set_dimension": (c_int64, [c_void_p]),
    "dh_sector_set_offsets": (c_int, [c_void_p, P_I64]),
    "dh_sector_set_get_states": (c_int, [c_void_p, P_U64]),
    "dh_sector_set_factors": (c_int, [c_void_p, P_D, P_I, c_int, c_double]),
    "dh_sector_set_launch": (c_int, [c_void_p, c_int, c_int]),
    "dh_sector_set_row_mode": (c_int, [c_void_p, c_int]),
    "dh_sector_set_blas_launch": (c_int, [c_void_p, c_int, c_int64]),
    "dh_sector_set_count_elements": (c_int, [c_void_p, P_I64]),
    "dh_sectors_hv": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_char_p]),
    "dh_sectors_block_dots": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_char_p, P_D]),
    "dh_sectors_block_axpy": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_char_p, P_D]),
    "dh_sectors_block_rmul": (c_int, [c_void_p, c_void_p, c_int, c_char_p, P_D]),
    "dh_momentum_sector_dimension": (c_int64, [c_int, c_int, c_int, c_int, c_int]),
    "dh_momentum_sector_set_create": (c_int, [c_void_p, PP, c_int, c_int, c_int, P_I, P_I, c_int]),
    "dh_sector_set_lattice_factors": (c_int, [c_void_p, c_int, P_I, P_I, P_D, c_double]),
    "dh_torus_translation_sector_set_create": (c_int, [c_void_p, PP, c_int, c_int, P_I, P_I, c_int]),
    "dh_sector_set_torus_data": (c_int, [c_void_p, c_void_p, P_U64]),
    "dh_block_dots": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int64, c_int, P_D]),
    "dh_block_dots_noalloc": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int64, c_int, c_void_p]),
    "dh_block_axpy": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int64, c_int, P_D]),
    "dh_block_rmul": (c_int, [c_void_p, c_void_p, c_int64, c_int, P_D]),
    "dh_block_update": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int64, c_int, P_D, P_D, P_D, P_D]),
    "dh_sectors_block_update": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_char_p, P_D, P_D, P_D, P_D]),
    "dh_block_get_column": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_int, c_void_p]),
    "dh_block_set_column": (c_int, [c_void_p, c_void_p, c_int64, c_int, c_int, c_void_p]),
    "dh_copy": (c_int, [c_void_p, c_void_p, c_void_p, c_int64]),
    "dh_axpby": (c_int, [c_void_p, c_double, c_void_p, c_double, c_void_p, c_int64]),
    "dh_fill_random": (c_int, [c_void_p, c_void_p, c_int64, c_uint64]),
}

HOST_SIGNATURES = {
    "dhh_last_error": (c_char_p, []),
    "dhh_last_solve_waves": (c_int, []),
    "dhh_sectors_launch": (c_int, [c_int, c_int, c_int]),
    "dhh_sectors_alone_rows": (c_int, [c_int64]),
    "dhh_sectors_autotune": (c_int, [c_void_p, c_int, c_int, P_I, c_int, P_D, c_int, P_I, c_int, c_double, c_int,
                                     c_uint64, P_D, P_I]),
    "dhh_arch_create": (c_int, [PP, c_int]),
    "dhh_arch_destroy": (c_int, [c_void_p]),
    "dhh_arch_context": (c_void_p, [c_void_p]),
    "dhh_sphere_problem_create": (c_int, [c_void_p, PP, c_int, c_int, c_int, P_D]),
    "dhh_sphere_nbody_problem_create": (c_int, [c_void_p, PP, c_int, c_int, c_int, c_int, P_I, c_int, P_I, P_D]),
    "dhh_sphere_l2_problem_create": (c_int, [c_void_p, PP, c_int, c_int, c_int, P_D, c_double]),
    "dhh_sphere_l2_factors": (c_int, [c_int, c_int, c_int, c_double, P_D, P_I, P_I, P_D]),
    "dhh_write_vector": (c_int, [c_void_p, c_char_p, c_void_p, c_int64, c_int]),
    "dhh_read_vector": (c_int, [c_void_p, c_char_p, c_void_p, c_int64, c_int]),
    "dhh_vector_file_dimension": (c_int, [c_char_p, c_int, P_I64]),
    "dhh_spin_problem_create": (c_int, [c_void_p, PP, c_int, c_int, P_D, P_D, c_int]),
    "dhh_problem_destroy": (c_int, [c_void_p]),
    "dhh_problem_basis": (c_void_p, [c_void_p]),
    "dhh_problem_hamiltonian": (c_void_p, [c_void_p]),
    "dhh_problem_channels": (c_int, [c_void_p, P_I, c_int]),
    "dhh_problem_multiple_multiply": (c_int, [c_void_p, PP, PP, c_int, c_int, c_int64, c_int64]),
    "dhh_lanczos": (c_int, [c_void_p, c_int, c_int, c_int, c_double, c_double, c_uint64, P_D, P_I, P_I, P_D,
                            c_void_p, c_int]),
    "dhh_lanczos_disk": (c_int, [c_void_p, c_char_p, c_int, c_int, c_int, c_int, c_double, c_double, c_uint64, P_D,
                                 P_I, P_I, P_D, c_void_p, c_int]),
    "dhh_sectors_solve": (c_int, [c_void_p, c_int, c_int, P_I, c_int, P_D, c_int, c_int, c_int, c_double, c_uint64,
                                  c_int, c_double, P_D, P_I, P_I, P_D]),
    "dhh_autotune": (c_int, [c_void_p, P_I, c_int, c_int, P_D, P_I]),
    "dhh_lattice_sectors_solve": (c_int, [c_void_p, c_int, c_int, c_int, c_double, c_double, c_double, c_double,
                                          c_double, c_double, c_int, P_I, P_I, c_int, c_int, c_int, c_int, c_double,
                                          c_uint64, c_double, P_D, P_I, P_I, P_D]),
    "dhh_checkerboard_lower_band": (c_int, [c_int, c_int, c_double, c_double, c_double, c_double, c_double,
                                            c_double, P_D, P_D]),
    "dhh_checkerboard_factors": (c_int, [c_int, c_int, c_double, c_double, c_double, c_double, c_double, c_double,
                                         c_int, P_I, P_I, P_I, P_D, P_I64]),
    "dhh_torus_factors": (c_int, [c_int, c_double, P_D, c_int, c_int, P_I, P_I, P_I, P_D, P_I64]),
    "dhh_torus_sectors_solve": (c_int, [c_void_p, c_int, c_int, c_double, P_D, c_int, c_int, P_I, c_int, c_int,
                                        c_int, c_int, c_double, c_uint64, c_double, P_D, P_I, P_I, P_D]),
    "dhh_torus_translation_sectors_solve": (c_int, [c_void_p, c_int, c_int, c_double, P_D, c_int, c_int, P_I, P_I,
                                                    c_int, c_int, c_int, c_int, c_double, c_uint64, c_double, P_D,
                                                    P_I, P_I, P_D]),
    "dhh_coulomb_pseudopotentials": (c_int, [c_int, c_int, P_D]),
    "dhh_clebsch_gordan": (c_double, [c_int, c_int, c_int, c_int, c_int]),
    "dhh_sphere_factors": (c_int, [c_int, P_D, P_D, P_I, P_I]),
    "dhh_sphere_channel_table": (c_int, [c_int, P_I, P_D, c_int]),
    "dhh_read_pseudopotentials": (c_int, [c_char_p, c_int, P_D]),
    "dhh_band_eigenvalues": (c_int, [c_int, c_int, P_D, P_D]),
}


def _bind(lib, sigs):
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


def gpu():
    global _gpu
    if _gpu is None:
        if not os.path.exists(LIB_GPU):
            raise RuntimeError("libdhgpu.so is not built: run `python -m diagham_amd.build`")
        _gpu = ctypes.CDLL(LIB_GPU, mode=ctypes.RTLD_GLOBAL)
        _bind(_gpu, GPU_SIGNATURES)
    return _gpu


def host():
    global _host
    if _host is None:
        gpu()
        if not os.path.exists(LIB_HOST):
            raise RuntimeError("libdiagham_host.so is not built: run `python -m diagham_amd.build`")
        _host = ctypes.CDLL(LIB_HOST, mode=ctypes.RTLD_GLOBAL)
        _bind(_host, HOST_SIGNATURES)
    return _host


class DHError(RuntimeError):
    pass


def dptr(a):
    return a.ctypes.data_as(P_D)


# ---------------------------------------------------------------------------- host-only helpers
def coulomb_pseudopotentials(nbr_flux, round_to_file=True):
    out = np.zeros(nbr_flux + 1)
    host().dhh_coulomb_pseudopotentials(nbr_flux, int(round_to_file), dptr(out))
    return out


def clebsch_gordan(j1, j2, m1, m2, j):
    return host().dhh_clebsch_gordan(j1, j2, m1, m2, j)


def sphere_factors(lz_max, pseudo):
    pseudo = np.ascontiguousarray(pseudo, dtype=np.float64)
    n_orb = lz_max + 1
    cap = n_orb * n_orb * n_orb
    f = np.zeros(cap)
    off = np.zeros(n_orb * (n_orb - 1) // 2 + 1, np.int32)
    n = ctypes.c_int(cap)
    rc = host().dhh_sphere_factors(lz_max, dptr(pseudo), dptr(f), off.ctypes.data_as(P_I), ctypes.byref(n))
    if rc != 0:
        raise DHError("dhh_sphere_factors failed: %d" % rc)
    return f[:n.value].copy(), off


def sphere_l2_factors(lz_max, n, total_lz, l2_factor=1.0):
    """factors of l2_factor * L^2 in the sphere factor layout, the pair offsets and the constant
    l2_factor * Lz (Lz + 1) of the sector (dhh_sphere_l2_factors)."""
    n_orb = lz_max + 1
    cap = n_orb * n_orb * n_orb
    f = np.zeros(cap)
    off = np.zeros(n_orb * (n_orb - 1) // 2 + 1, np.int32)
    nf = ctypes.c_int(cap)
    shift = ctypes.c_double()
    rc = host().dhh_sphere_l2_factors(lz_max, n, total_lz, l2_factor, dptr(f), off.ctypes.data_as(P_I),
                                      ctypes.byref(nf), ctypes.byref(shift))
    if rc != 0:
        raise DHError("dhh_sphere_l2_factors: " + host().dhh_last_error().decode())
    return f[:nf.value].copy(), off, shift.value


def sphere_channel_table(lz_max):
    n_orb = lz_max + 1
    npairs = n_orb * (n_orb - 1) // 2
    cap = npairs * (lz_max + 1)
    mom = np.zeros(lz_max + 1, np.int32)
    tab = np.zeros(cap)
    nc = host().dhh_sphere_channel_table(lz_max, mom.ctypes.data_as(P_I), dptr(tab), cap)
    return mom[:nc].copy(), tab[:npairs * nc].reshape(npairs, nc).copy()


def band_eigenvalues(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    ev = np.zeros(n)
    rc = host().dhh_band_eigenvalues(n, b, dptr(a), dptr(ev))
    if rc != 0:
        raise DHError(host().dhh_last_error().decode())
    return ev


def fermion_sphere_dimension(n, total_lz, lz_max):
    return gpu().dh_fermion_sphere_dimension(n, total_lz, lz_max)


# ---------------------------------------------------------------------------- GPU objects
_live_archs = weakref.WeakSet()


@atexit.register
def _close_live_archs():
    """Close every open Arch (and what it tracks) while everything is still reachable, before the
    interpreter's final collections run finalizers in no defined order."""
    for a in list(_live_archs):
        try:
            a.close()
        except Exception:
            pass


class Arch:
    """GPUArchitecture: one device context.

    The native layers count references: every problem, sector set and buffer keeps its context
    alive until it is released, so objects may be closed or collected in any order, before or
    after the Arch.  close() releases the tracked objects first and drops the handles it held;
    calls on a closed Arch then pass null handles, which the C ABI refuses."""

    def __init__(self, device=0):
        self._dependents = weakref.WeakSet()
        h = c_void_p()
        rc = host().dhh_arch_create(ctypes.byref(h), device)
        if rc != 0:
            raise DHError("dhh_arch_create failed (%d): %s" % (rc, host().dhh_last_error().decode()))
        self.h = h
        self.ctx = host().dhh_arch_context(h)
        _live_archs.add(self)

    def track(self, obj):
        """Register an object whose native handle lives in this context (it needs `close()`)."""
        self._dependents.add(obj)
        return obj

    def close(self):
        """Release every live buffer, problem and sector set of this context, then the context."""
        if getattr(self, "h", None):
            for obj in list(getattr(self, "_dependents", ())):
                try:
                    obj.close()
                except Exception:
                    pass
            host().dhh_arch_destroy(self.h)
            self.h = None
            self.ctx = None

    def __del__(self):
        self.close()

    def check(self, rc, what):
        if rc != 0:
            raise DHError("%s failed (%d): %s" % (what, rc, gpu().dh_last_error(self.ctx).decode()))

    def device_info(self):
        n = ctypes.c_int()
        name = ctypes.create_string_buffer(64)
        self.check(gpu().dh_device_info(self.ctx, ctypes.byref(n), name), "dh_device_info")
        return n.value, name.value.decode()

    def sync(self):
        self.check(gpu().dh_synchronize(self.ctx), "dh_synchronize")

    def memory_peak(self, reset=False):
        """highest device memory in use since creation or the last reset (dh_memory_peak)"""
        n = ctypes.c_size_t()
        self.check(gpu().dh_memory_peak(self.ctx, ctypes.byref(n), int(reset)), "dh_memory_peak")
        return n.value


class DeviceBuffer:
    """Device memory from dh_malloc holding float64 (or raw bytes)."""

    def __init__(self, arch, n, dtype=np.float64):
        self.arch = arch
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        p = c_void_p()
        arch.check(gpu().dh_malloc(arch.ctx, ctypes.byref(p), max(self.n, 1) * self.dtype.itemsize), "dh_malloc")
        self.ptr = p.value
        # the context this buffer was allocated in: the buffer keeps it alive until dh_free
        self._ctx = arch.ctx
        arch.track(self)

    @classmethod
    def from_host(cls, arch, a):
        a = np.ascontiguousarray(a)
        b = cls(arch, a.size, a.dtype)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        self.arch.check(gpu().dh_memcpy_h2d(self.arch.ctx, self.ptr, a.ctypes.data, a.nbytes), "h2d")

    def download(self):
        out = np.empty(self.n, self.dtype)
        self.arch.check(gpu().dh_memcpy_d2h(self.arch.ctx, out.ctypes.data, self.ptr, out.nbytes), "d2h")
        return out

    def zero(self):
        self.arch.check(gpu().dh_memset(self.arch.ctx, self.ptr, 0, self.n * self.dtype.itemsize), "memset")

    def free(self):
        if getattr(self, "ptr", None):
            gpu().dh_free(self._ctx, self.ptr)
            self.ptr = None

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Problem:
    """A sector: Hilbert space + Hamiltonian built by the C++ host layer."""

    def __init__(self, arch, handle, kind):
        self.arch = arch
        self.h = handle
        arch.track(self)
        self.kind = kind
        self.ham = host().dhh_problem_hamiltonian(handle)
        self.basis = host().dhh_problem_basis(handle)
        self.dim = gpu().dh_basis_dimension(self.basis)

    @classmethod
    def sphere(cls, arch, n, total_lz, lz_max, pseudo):
        pseudo = np.ascontiguousarray(pseudo, dtype=np.float64)
        assert pseudo.size >= lz_max + 1
        h = c_void_p()
        rc = host().dhh_sphere_problem_create(arch.h, ctypes.byref(h), n, total_lz, lz_max, dptr(pseudo))
        if rc != 0:
            raise DHError("sphere problem: " + host().dhh_last_error().decode())
        p = cls(arch, h, "sphere")
        p.n, p.total_lz, p.lz_max = n, total_lz, lz_max
        return p

    @classmethod
    def sphere_l2(cls, arch, n, total_lz, lz_max, l2_factor=1.0, pseudo=None):
        """l2_factor * L^2 (plus the pseudopotential interaction if given): dhh_sphere_l2_problem_create."""
        pp = None if pseudo is None else np.ascontiguousarray(pseudo, dtype=np.float64)
        h = c_void_p()
        rc = host().dhh_sphere_l2_problem_create(arch.h, ctypes.byref(h), n, total_lz, lz_max,
                                                 None if pp is None else dptr(pp), l2_factor)
        if rc != 0:
            raise DHError("L^2 problem: " + host().dhh_last_error().decode())
        p = cls(arch, h, "sphere_l2")
        p.n, p.total_lz, p.lz_max = n, total_lz, lz_max
        return p

    @classmethod
    def sphere_nbody(cls, arch, n, total_lz, lz_max, nbody, sum_count, tuples, factors):
        """Three-body term with explicit per-sum factor matrices (dhh_sphere_nbody_problem_create)."""
        sum_count = np.ascontiguousarray(sum_count, dtype=np.int32)
        tuples = np.ascontiguousarray(tuples, dtype=np.int32)
        factors = np.ascontiguousarray(factors, dtype=np.float64)
        h = c_void_p()
        rc = host().dhh_sphere_nbody_problem_create(arch.h, ctypes.byref(h), n, total_lz, lz_max, nbody,
                                                    sum_count.ctypes.data_as(P_I), sum_count.size,
                                                    tuples.ctypes.data_as(P_I), dptr(factors))
        if rc != 0:
            raise DHError("n-body problem: " + host().dhh_last_error().decode())
        p = cls(arch, h, "sphere_nbody")
        p.n, p.total_lz, p.lz_max = n, total_lz, lz_max
        return p

    @classmethod
    def spin_chain(cls, arch, length, sz, j=1.0, jz=1.0, periodic=True):
        J = np.full(length, j, np.float64) if np.isscalar(j) else np.ascontiguousarray(j, np.float64)
        Jz = np.full(length, jz, np.float64) if np.isscalar(jz) else np.ascontiguousarray(jz, np.float64)
        h = c_void_p()
        rc = host().dhh_spin_problem_create(arch.h, ctypes.byref(h), length, sz, dptr(J), dptr(Jz), int(periodic))
        if rc != 0:
            raise DHError("spin problem: " + host().dhh_last_error().decode())
        p = cls(arch, h, "spin")
        p.length = length
        return p

    def close(self):
        if getattr(self, "h", None):
            host().dhh_problem_destroy(self.h)
            self.h = None
            self.ham = None
            self.basis = None

    def __del__(self):
        self.close()

    # -- basis
    def states(self):
        out = np.zeros(self.dim, np.uint64)
        self.arch.check(gpu().dh_basis_get_states(self.basis, out.ctypes.data_as(P_U64)), "get_states")
        return out

    def find_state_index(self, states, mode=INDEX_RANK):
        states = np.ascontiguousarray(states, dtype=np.uint64)
        out = np.zeros(states.size, np.int64)
        self.arch.check(gpu().dh_basis_find_state_index(self.basis, states.ctypes.data_as(P_U64), states.size,
                                                        mode, out.ctypes.data_as(P_I64)), "find_state_index")
        return out

    # -- Hamiltonian
    def set_shift(self, shift):
        self.arch.check(gpu().dh_hamiltonian_set_shift(self.ham, shift), "set_shift")

    def set_launch(self, lanes_per_row=1, block_size=256, index_mode=INDEX_RANK, lds_tables=1):
        self.arch.check(gpu().dh_hamiltonian_set_launch(self.ham, lanes_per_row, block_size, index_mode,
                                                        lds_tables), "set_launch")

    def hv_device(self, d_x, d_y, k=1, add=False, first=0, count=-1):
        self.arch.check(gpu().dh_hv(self.ham, d_x, d_y, k, int(add), first, count), "dh_hv")

    def hv_vectors_device(self, d_xs, d_ys, add=False, first=0, count=-1, via_host=False):
        """H applied to len(d_xs) separate device vectors (dh_hv_vectors; via_host: through the
        host layer's MultipleVectorHamiltonianMultiplyOperation, dhh_problem_multiple_multiply)"""
        n = len(d_xs)
        assert len(d_ys) == n
        xs = (c_void_p * max(n, 1))(*d_xs)
        ys = (c_void_p * max(n, 1))(*d_ys)
        if via_host:
            rc = host().dhh_problem_multiple_multiply(self.h, xs, ys, n, int(add), first, count)
            if rc != 0:
                raise DHError("dhh_problem_multiple_multiply: " + host().dhh_last_error().decode())
        else:
            self.arch.check(gpu().dh_hv_vectors(self.ham, xs, ys, n, int(add), first, count), "dh_hv_vectors")

    def hv(self, x, k=1):
        """H x for host arrays (x: dim*k, [row][k] layout); convenience for tests."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        dx = DeviceBuffer.from_host(self.arch, x)
        dy = DeviceBuffer(self.arch, x.size)
        self.hv_device(dx.ptr, dy.ptr, k)
        return dy.download()

    def count_elements(self):
        n = ctypes.c_int64()
        self.arch.check(gpu().dh_hamiltonian_count_elements(self.ham, ctypes.byref(n)), "count_elements")
        return n.value

    def row_elements(self, row, max_out=100000):
        t = np.zeros(max_out, np.int64)
        c = np.zeros(max_out)
        n = ctypes.c_int64()
        self.arch.check(gpu().dh_hamiltonian_row_elements(self.ham, row, t.ctypes.data_as(P_I64), dptr(c),
                                                          max_out, ctypes.byref(n)), "row_elements")
        assert n.value <= max_out
        return t[:n.value].copy(), c[:n.value].copy()

    def channels(self):
        mom = np.zeros(64, np.int32)
        n = host().dhh_problem_channels(self.h, mom.ctypes.data_as(P_I), 64)
        return mom[:n].copy()

    def adjoint_device(self, d_psi, d_phi, d_hpsi=None):
        nc = len(self.channels())
        g = np.zeros(nc)
        self.arch.check(gpu().dh_sphere_adjoint(self.ham, d_psi, d_phi, dptr(g), d_hpsi), "dh_sphere_adjoint")
        return g

    def precompute(self, max_bytes=1 << 40):
        """Store every matrix element in device memory (dh_hamiltonian_precompute); returns the
        bytes used.  Raises if they would exceed max_bytes (the products then stay on the fly)."""
        n = ctypes.c_int64()
        rc = gpu().dh_hamiltonian_precompute(self.ham, max_bytes, ctypes.byref(n))
        if rc != 0:
            raise DHError("dh_hamiltonian_precompute (%d bytes needed): %s" % (
                n.value, gpu().dh_last_error(self.arch.ctx).decode()))
        return n.value

    def release_precomputed(self):
        self.arch.check(gpu().dh_hamiltonian_release_precomputed(self.ham), "release_precomputed")

    def autotune(self, ks=(1, 2, 4, 8, 16), repeats=5):
        """Time every launch shape per block size; returns (table rows, best k) and leaves the
        fastest launch of the best k configured (see dhh_autotune)."""
        ks = np.ascontiguousarray(ks, dtype=np.int32)
        table = np.zeros(6 * ks.size)
        best = ctypes.c_int()
        rc = host().dhh_autotune(self.h, ks.ctypes.data_as(P_I), ks.size, repeats, dptr(table), ctypes.byref(best))
        if rc != 0:
            raise DHError("dhh_autotune: " + host().dhh_last_error().decode())
        rows = [dict(zip(("k", "lanes_per_row", "block_size", "index_mode", "ms", "elements_per_s"), r))
                for r in table.reshape(-1, 6)]
        return rows, best.value

    # -- Lanczos
    def lanczos(self, nbr_eigen, block_size=1, max_iter=3000, precision=1e-14, seed=0x5EED0000,
                d_vectors=None, nbr_vectors=0, allow_unconverged=False, residual=0.0, disk=None, resume=False):
        """disk: directory of the disk storage (dhh_lanczos_disk); resume continues the run stored
        there."""
        ev = np.zeros(nbr_eigen)
        prods, blocks, secs = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        if disk is not None:
            rc = host().dhh_lanczos_disk(self.h, os.fsencode(disk), int(resume), nbr_eigen, block_size, max_iter,
                                         precision, residual, seed, dptr(ev), ctypes.byref(prods),
                                         ctypes.byref(blocks), ctypes.byref(secs), d_vectors, nbr_vectors)
        else:
            if resume:
                raise ValueError("resume needs the disk storage directory")
            rc = host().dhh_lanczos(self.h, nbr_eigen, block_size, max_iter, precision, residual, seed, dptr(ev),
                                    ctypes.byref(prods), ctypes.byref(blocks), ctypes.byref(secs), d_vectors,
                                    nbr_vectors)
        if rc != 0 and not (allow_unconverged and rc == DH_E_STATE):
            raise DHError("dhh_lanczos failed (%d): %s" % (rc, host().dhh_last_error().decode()))
        return {"eigenvalues": ev, "products": prods.value, "blocks": blocks.value, "seconds": secs.value,
                "converged": rc == 0}


class SphereSectorSet:
    """dh_sector_set of fermions on the sphere: the listed total_lz sectors concatenated, with the
    interaction factors of the pseudopotentials."""

    def __init__(self, arch, n, lz_max, total_lz, pseudo):
        self.arch = arch
        lz = np.ascontiguousarray(total_lz, dtype=np.int32)
        h = ctypes.c_void_p()
        arch.check(gpu().dh_sector_set_create(arch.ctx, ctypes.byref(h), n, lz_max, lz.ctypes.data_as(P_I), lz.size),
                   "dh_sector_set_create")
        self.h = h
        arch.track(self)
        self.nbr_sectors = lz.size
        self.offsets = np.zeros(lz.size + 1, np.int64)
        gpu().dh_sector_set_offsets(h, self.offsets.ctypes.data_as(P_I64))
        self.total = int(self.offsets[-1])
        f, po = sphere_factors(lz_max, pseudo)
        arch.check(gpu().dh_sector_set_factors(h, dptr(f), po.ctypes.data_as(P_I), f.size, 0.0), "factors")

    def set_launch(self, lanes_per_row=1, block_size=256):
        self.arch.check(gpu().dh_sector_set_launch(self.h, lanes_per_row, block_size), "dh_sector_set_launch")

    def set_row_mode(self, mode):
        self.arch.check(gpu().dh_sector_set_row_mode(self.h, mode), "dh_sector_set_row_mode")

    def set_blas_launch(self, axpy_groups_per_cu=16, dots_rows_per_tile=0):
        self.arch.check(gpu().dh_sector_set_blas_launch(self.h, axpy_groups_per_cu, dots_rows_per_tile),
                        "dh_sector_set_blas_launch")

    def count_elements(self):
        counts = np.zeros(self.nbr_sectors, np.int64)
        self.arch.check(gpu().dh_sector_set_count_elements(self.h, counts.ctypes.data_as(P_I64)), "count")
        return counts

    def hv_device(self, d_x, d_y, k=1, add=False, active=None):
        am = None if active is None else bytes(np.asarray(active, np.uint8))
        self.arch.check(gpu().dh_sectors_hv(self.h, d_x, d_y, k, int(add), am), "dh_sectors_hv")

    def close(self):
        if getattr(self, "h", None):
            gpu().dh_sector_set_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def checkerboard_factors(nx, ny, t1=1.0, t2=0.5, tpp=-0.5, mus=0.0, u_potential=1.0, v_potential=0.0,
                         flat_band=True):
    """dhh_checkerboard_factors: (sum_pair_count, sum_pairs, complex factors interleaved) of the
    projected checkerboard model in the momentum sector set layout"""
    npairs = nx * ny * (nx * ny - 1) // 2
    cnt = np.zeros(nx * ny, np.int32)
    pairs = np.zeros(2 * npairs, np.int32)
    nf = ctypes.c_int64(npairs * npairs)
    npr = ctypes.c_int(npairs)
    f = np.zeros(2 * npairs * npairs)
    rc = host().dhh_checkerboard_factors(nx, ny, t1, t2, tpp, mus, u_potential, v_potential, int(flat_band),
                                         cnt.ctypes.data_as(P_I), pairs.ctypes.data_as(P_I), ctypes.byref(npr),
                                         dptr(f), ctypes.byref(nf))
    if rc != 0:
        raise DHError("dhh_checkerboard_factors failed (%d)" % rc)
    return cnt, pairs[:2 * npr.value], f[:2 * nf.value]


class MomentumSectorSet:
    """dh_sector_set of fermions in momentum space (the (kx[i], ky[i]) sectors concatenated) with
    the projected checkerboard factors; vectors are complex, (re, im) interleaved"""

    def __init__(self, arch, n, nx, ny, kx, ky, **model):
        self.arch = arch
        kxa = np.ascontiguousarray(kx, dtype=np.int32)
        kya = np.ascontiguousarray(ky, dtype=np.int32)
        h = ctypes.c_void_p()
        arch.check(gpu().dh_momentum_sector_set_create(arch.ctx, ctypes.byref(h), n, nx, ny, kxa.ctypes.data_as(P_I),
                                                       kya.ctypes.data_as(P_I), kxa.size), "momentum sector set")
        self.h = h
        arch.track(self)
        self.nbr_sectors = kxa.size
        self.offsets = np.zeros(kxa.size + 1, np.int64)
        gpu().dh_sector_set_offsets(h, self.offsets.ctypes.data_as(P_I64))
        self.total = int(self.offsets[-1])
        cnt, pairs, f = checkerboard_factors(nx, ny, **model)
        arch.check(gpu().dh_sector_set_lattice_factors(h, nx * ny, cnt.ctypes.data_as(P_I), pairs.ctypes.data_as(P_I),
                                                       dptr(f), 0.0), "lattice factors")

    def set_launch(self, lanes_per_row=1, block_size=256):
        self.arch.check(gpu().dh_sector_set_launch(self.h, lanes_per_row, block_size), "dh_sector_set_launch")

    def hv_device(self, d_x, d_y, k=1, add=False, active=None):
        am = None if active is None else bytes(np.asarray(active, np.uint8))
        self.arch.check(gpu().dh_sectors_hv(self.h, d_x, d_y, k, int(add), am), "dh_sectors_hv")

    def count_elements(self):
        counts = np.zeros(self.nbr_sectors, np.int64)
        self.arch.check(gpu().dh_sector_set_count_elements(self.h, counts.ctypes.data_as(P_I64)), "count")
        return counts

    def close(self):
        if getattr(self, "h", None):
            gpu().dh_sector_set_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lattice_set(arch, n, nx, ny, kx, ky, **model):
    return MomentumSectorSet(arch, n, nx, ny, kx, ky, **model)


def sphere_sectors(n, lz_max, positive_only=True):
    """total_lz values (twice Lz) of the non-empty sectors of n fermions at lz_max; Lz >= 0 only
    by default, as in the Lz loop of FQHESphereFermionsTwoBodyGeneric.cc:270."""
    top = n * (lz_max - n + 1)
    low = top % 2 if positive_only else -top
    return list(range(low, top + 1, 2))


def sectors_solve(arch, n, lz_max, total_lz, pseudo, nbr_eigen, block_size=1, batched=True, max_iter=3000,
                  precision=1e-14, seed=0x5EED0000, memory_gb=0.0, allow_unconverged=False):
    """dhh_sectors_solve: the nbr_eigen lowest eigenvalues of every sector in total_lz."""
    lz = np.ascontiguousarray(total_lz, dtype=np.int32)
    pseudo = np.ascontiguousarray(pseudo, dtype=np.float64)
    ns = lz.size
    ev = np.zeros(ns * nbr_eigen)
    prods = np.zeros(ns, np.int32)
    conv = np.zeros(ns, np.int32)
    secs = ctypes.c_double()
    rc = host().dhh_sectors_solve(arch.h, n, lz_max, lz.ctypes.data_as(P_I), ns, dptr(pseudo), nbr_eigen,
                                  block_size, max_iter, precision, seed, int(bool(batched)), memory_gb, dptr(ev),
                                  prods.ctypes.data_as(P_I), conv.ctypes.data_as(P_I), ctypes.byref(secs))
    if rc != 0 and not (allow_unconverged and rc == DH_E_STATE):
        raise DHError("dhh_sectors_solve failed (%d): %s" % (rc, host().dhh_last_error().decode()))
    return {"eigenvalues": ev.reshape(ns, nbr_eigen), "products": prods, "converged": conv.astype(bool),
            "seconds": secs.value, "waves": host().dhh_last_solve_waves()}


def lattice_sectors_solve(arch, n, nx, ny, kx, ky, nbr_eigen, block_size=1, t1=1.0, t2=0.5, tpp=-0.5, mus=0.0,
                          u_potential=1.0, v_potential=0.0, flat_band=True, max_iter=3000, precision=1e-14,
                          seed=0x5EED0000, memory_gb=0.0, allow_unconverged=False):
    """dhh_lattice_sectors_solve: the nbr_eigen lowest eigenvalues of every (kx[i], ky[i]) sector of
    the projected checkerboard model, all sectors advanced together"""
    kxa = np.ascontiguousarray(kx, dtype=np.int32)
    kya = np.ascontiguousarray(ky, dtype=np.int32)
    assert kxa.size == kya.size
    ns = kxa.size
    ev = np.zeros(ns * nbr_eigen)
    prods = np.zeros(ns, np.int32)
    conv = np.zeros(ns, np.int32)
    secs = ctypes.c_double()
    rc = host().dhh_lattice_sectors_solve(arch.h, n, nx, ny, t1, t2, tpp, mus, u_potential, v_potential,
                                          int(flat_band), kxa.ctypes.data_as(P_I), kya.ctypes.data_as(P_I), ns,
                                          nbr_eigen, block_size, max_iter, precision, seed, memory_gb, dptr(ev),
                                          prods.ctypes.data_as(P_I), conv.ctypes.data_as(P_I), ctypes.byref(secs))
    if rc != 0 and not (allow_unconverged and rc == DH_E_STATE):
        raise DHError("dhh_lattice_sectors_solve failed (%d): %s" % (rc, host().dhh_last_error().decode()))
    return {"eigenvalues": ev.reshape(ns, nbr_eigen), "products": prods, "converged": conv.astype(bool),
            "seconds": secs.value, "waves": host().dhh_last_solve_waves()}


def momentum_sector_dims(arch, n, nx, ny, kx, ky):
    """dimensions of the (kx[i], ky[i]) sectors of n fermions in momentum space on nx x ny sites"""
    kxa = np.ascontiguousarray(kx, dtype=np.int32)
    kya = np.ascontiguousarray(ky, dtype=np.int32)
    h = ctypes.c_void_p()
    arch.check(gpu().dh_momentum_sector_set_create(arch.ctx, ctypes.byref(h), n, nx, ny, kxa.ctypes.data_as(P_I),
                                                   kya.ctypes.data_as(P_I), kxa.size), "momentum sector set")
    off = np.zeros(kxa.size + 1, np.int64)
    gpu().dh_sector_set_offsets(h, off.ctypes.data_as(P_I64))
    gpu().dh_sector_set_destroy(h)
    return np.diff(off).tolist()


def sectors_launch(lanes_per_row=0, block_size=256, row_mode=INDEX_RANK):
    """launch shape of later batched sector solves of this thread (dhh_sectors_launch; 0: built-in)"""
    rc = host().dhh_sectors_launch(lanes_per_row, block_size, row_mode)
    if rc != 0:
        raise DHError("dhh_sectors_launch: " + host().dhh_last_error().decode())


def sectors_autotune(arch, n, lz_max, total_lz, pseudo, nbr_eigen, ks=(1, 2, 4, 8, 16), fraction=0.05,
                     max_iter=3000, seed=0x5EED0000):
    """dhh_sectors_autotune: per k the fastest product launch on calibration sectors, then the
    time of a calibration solve with it; returns (rows of the table, best shape dict)"""
    lz = np.ascontiguousarray(total_lz, dtype=np.int32)
    ka = np.ascontiguousarray(ks, dtype=np.int32)
    pp = np.ascontiguousarray(pseudo, dtype=np.float64)
    table = np.zeros(7 * ka.size)
    best = np.zeros(4, np.int32)
    rc = host().dhh_sectors_autotune(arch.h, n, lz_max, lz.ctypes.data_as(P_I), lz.size, dptr(pp), nbr_eigen,
                                     ka.ctypes.data_as(P_I), ka.size, fraction, max_iter, seed, dptr(table),
                                     best.ctypes.data_as(P_I))
    if rc != 0:
        raise DHError("dhh_sectors_autotune: " + host().dhh_last_error().decode())
    keys = ("k", "lanes_per_row", "block_size", "row_mode", "product_ms", "calibration_seconds",
            "calibration_products")
    rows = [dict(zip(keys, [float(v) for v in r])) for r in table.reshape(-1, 7)]
    return rows, dict(zip(("k", "lanes_per_row", "block_size", "row_mode"), [int(v) for v in best]))


def torus_factors(nbr_flux, ratio=1.0, pseudo=(), coulomb_level=-1):
    """dhh_torus_factors: (sum_pair_count, sum_pairs, complex factors interleaved) of the torus
    Hamiltonian in the momentum sector set layout (Nx = 1)."""
    pp = np.ascontiguousarray(pseudo, dtype=np.float64)
    npairs = nbr_flux * (nbr_flux - 1) // 2
    cnt = np.zeros(nbr_flux, np.int32)
    pairs = np.zeros(2 * npairs, np.int32)
    f = np.zeros(2 * npairs * npairs)
    npr, nf = ctypes.c_int(npairs), ctypes.c_int64(npairs * npairs)
    rc = host().dhh_torus_factors(nbr_flux, ratio, dptr(pp) if pp.size else None, pp.size, coulomb_level,
                                  cnt.ctypes.data_as(P_I), pairs.ctypes.data_as(P_I), ctypes.byref(npr), dptr(f),
                                  ctypes.byref(nf))
    if rc != 0:
        raise DHError("dhh_torus_factors failed (%d): %s" % (rc, host().dhh_last_error().decode()))
    return cnt, pairs[:2 * npr.value], f[:2 * nf.value]


def torus_sectors_solve(arch, n, nbr_flux, ky, nbr_eigen, ratio=1.0, pseudo=(), coulomb_level=-1, block_size=1,
                        max_iter=3000, precision=1e-14, seed=0x5EED0000, memory_gb=0.0, allow_unconverged=False):
    """dhh_torus_sectors_solve: the nbr_eigen lowest eigenvalues of every Ky sector in ky."""
    k = np.ascontiguousarray(ky, dtype=np.int32)
    pp = np.ascontiguousarray(pseudo, dtype=np.float64)
    ns = k.size
    ev = np.zeros(ns * nbr_eigen)
    prods = np.zeros(ns, np.int32)
    conv = np.zeros(ns, np.int32)
    secs = ctypes.c_double()
    rc = host().dhh_torus_sectors_solve(arch.h, n, nbr_flux, ratio, dptr(pp) if pp.size else None, pp.size,
                                        coulomb_level, k.ctypes.data_as(P_I), ns, nbr_eigen, block_size, max_iter,
                                        precision, seed, memory_gb, dptr(ev), prods.ctypes.data_as(P_I),
                                        conv.ctypes.data_as(P_I), ctypes.byref(secs))
    if rc != 0 and not (allow_unconverged and rc == DH_E_STATE):
        raise DHError("dhh_torus_sectors_solve failed (%d): %s" % (rc, host().dhh_last_error().decode()))
    return {"eigenvalues": ev.reshape(ns, nbr_eigen), "products": prods, "converged": conv.astype(bool),
            "seconds": secs.value, "waves": host().dhh_last_solve_waves()}


def torus_translation_sectors_solve(arch, n, nbr_flux, kx, ky, nbr_eigen, ratio=1.0, pseudo=(), coulomb_level=-1,
                                    block_size=1, max_iter=3000, precision=1e-14, seed=0x5EED0000, memory_gb=0.0,
                                    allow_unconverged=False):
    """dhh_torus_translation_sectors_solve: the nbr_eigen lowest eigenvalues of every (kx[i], ky[i])
    sector of the torus with magnetic translations."""
    kxa = np.ascontiguousarray(kx, dtype=np.int32)
    kya = np.ascontiguousarray(ky, dtype=np.int32)
    assert kxa.size == kya.size
    pp = np.ascontiguousarray(pseudo, dtype=np.float64)
    ns = kxa.size
    ev = np.zeros(ns * nbr_eigen)
    prods = np.zeros(ns, np.int32)
    conv = np.zeros(ns, np.int32)
    secs = ctypes.c_double()
    rc = host().dhh_torus_translation_sectors_solve(
        arch.h, n, nbr_flux, ratio, dptr(pp) if pp.size else None, pp.size, coulomb_level, kxa.ctypes.data_as(P_I),
        kya.ctypes.data_as(P_I), ns, nbr_eigen, block_size, max_iter, precision, seed, memory_gb, dptr(ev),
        prods.ctypes.data_as(P_I), conv.ctypes.data_as(P_I), ctypes.byref(secs))
    if rc != 0 and not (allow_unconverged and rc == DH_E_STATE):
        raise DHError("dhh_torus_translation_sectors_solve failed (%d): %s" % (rc, host().dhh_last_error().decode()))
    return {"eigenvalues": ev.reshape(ns, nbr_eigen), "products": prods, "converged": conv.astype(bool),
            "seconds": secs.value, "waves": host().dhh_last_solve_waves()}
