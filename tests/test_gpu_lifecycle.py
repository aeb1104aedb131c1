"""Handle lifetimes on the device: the context outlives every object built on it.

A context is released by dh_destroy and by each of its bases, Hamiltonians, sector sets and
dh_malloc buffers; it is freed with the last of them.  The subprocess test builds those objects
around an Arch, drops them in a reference cycle and leaves another set open at exit, with the
exit hook that closes open Archs removed, so the finalizers run in whatever order the collector
picks.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu

SCRIPT = r"""
import atexit, gc, sys
sys.path.insert(0, %r)
import numpy as np
from diagham_amd import capi
atexit.unregister(capi._close_live_archs)

def build():
    a = capi.Arch(0)
    V = capi.coulomb_pseudopotentials(15, True)
    p = capi.Problem.sphere(a, 6, 0, 15, V)
    s = capi.SphereSectorSet(a, 6, 15, [0, 2], V)
    x = capi.DeviceBuffer.from_host(a, np.ones(p.dim))
    y = capi.DeviceBuffer(a, p.dim)
    p.hv_device(x.ptr, y.ptr)
    a.cycle = (p, s, x, y)
    return a

build()
gc.collect()
kept = build()   # still open when the interpreter exits
print("built", flush=True)
"""


def test_objects_released_in_any_order_at_exit():
    print("lifecycle subprocess: sphere N=6 2S=15 problem, sector set, two buffers, twice", flush=True)
    r = subprocess.run([sys.executable, "-c", SCRIPT % ROOT], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert "built" in r.stdout


def test_buffer_and_basis_outlive_dh_destroy():
    from diagham_amd import capi
    g = capi.gpu()
    ctx = ctypes.c_void_p()
    assert g.dh_create(ctypes.byref(ctx), 0) == 0
    buf = ctypes.c_void_p()
    assert g.dh_malloc(ctx, ctypes.byref(buf), 1024) == 0
    basis = ctypes.c_void_p()
    assert g.dh_fermion_sphere_create(ctx, ctypes.byref(basis), 4, 0, 8) == 0
    assert g.dh_destroy(ctx) == 0
    # the buffer and the basis hold the context: it is still there, closed for new objects
    assert g.dh_destroy(ctx) == capi.DH_E_STATE
    other = ctypes.c_void_p()
    assert g.dh_fermion_sphere_create(ctx, ctypes.byref(other), 4, 0, 8) == capi.DH_E_STATE
    states = np.zeros(g.dh_basis_dimension(basis), np.uint64)
    assert g.dh_basis_get_states(basis, states.ctypes.data_as(capi.P_U64)) == 0
    assert states.size == capi.fermion_sphere_dimension(4, 0, 8) == 12
    assert np.all(np.diff(states.astype(np.int64)) < 0)
    assert g.dh_free(ctx, buf) == 0
    assert g.dh_basis_destroy(basis) == 0   # the last reference: the context goes with it


def test_dh_free_refuses_foreign_pointers(arch):
    from diagham_amd import capi
    g = capi.gpu()
    b = capi.DeviceBuffer(arch, 8)
    inner = ctypes.c_void_p(b.ptr + 8)
    assert g.dh_free(arch.ctx, inner) == capi.DH_E_ARG
    b.free()
    assert g.dh_free(arch.ctx, ctypes.c_void_p(b.ptr)) == 0  # None: nothing to free
