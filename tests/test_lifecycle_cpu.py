"""Handle lifetimes across the C ABI, on the CPU.

The native libraries count references: a context lives until dh_destroy and every basis,
Hamiltonian, sector set and dh_malloc buffer of it are released; an architecture of the host layer
lives until dhh_arch_destroy and every problem built on it are released.  These tests drive
diagham_amd.capi against a stand-in of the two libraries that keeps exactly that contract and
records every call that reaches a freed context.  The stand-in can also free a context at
dhh_arch_destroy regardless of what still uses it: the objects of an Arch collected together
with it then reach the freed context, the fault seen at interpreter exit before the counting.
"""
import ctypes
import gc
import types

import pytest

from diagham_amd import capi


class FakeNative:
    """Both libraries' lifetime rules; `counting=False` frees a context at dhh_arch_destroy."""

    def __init__(self, counting=True):
        self.counting = counting
        self.next_id = 1000
        self.ctx = {}      # id -> {"refs": n, "freed": bool}
        self.arch = {}     # id -> {"refs": n, "ctx": id, "closed": bool}
        self.bufs = {}     # id -> ctx id
        self.problems = {}  # id -> arch id
        self.faults = []
        self.log = []

    def _new(self):
        self.next_id += 8
        return self.next_id

    def _use_ctx(self, c, what):
        if c is None or c not in self.ctx:
            self.faults.append((what, "unknown context %r" % (c,)))
        elif self.ctx[c]["freed"]:
            self.faults.append((what, "context used after free"))
        self.log.append(what)

    def _release_ctx(self, c):
        self.ctx[c]["refs"] -= 1
        if self.ctx[c]["refs"] == 0:
            self.ctx[c]["freed"] = True

    def _release_arch(self, a):
        self.arch[a]["refs"] -= 1
        if self.arch[a]["refs"] == 0:
            self._release_ctx(self.arch[a]["ctx"])

    # -- host layer
    def dhh_arch_create(self, out, device):
        a, c = self._new(), self._new()
        self.ctx[c] = {"refs": 1, "freed": False}
        self.arch[a] = {"refs": 1, "ctx": c, "closed": False}
        out._obj.value = a
        self.log.append("arch_create")
        return 0

    def dhh_arch_context(self, h):
        return self.arch[h.value]["ctx"]

    def dhh_arch_destroy(self, h):
        a = h.value
        self._use_ctx(self.arch[a]["ctx"], "arch_destroy")
        if self.arch[a]["closed"]:
            self.faults.append(("arch_destroy", "destroyed twice"))
            return capi.DH_E_STATE
        self.arch[a]["closed"] = True
        if self.counting:
            self._release_arch(a)
        else:
            self.ctx[self.arch[a]["ctx"]]["freed"] = True
        return 0

    def new_problem(self, arch_h):
        p = self._new()
        self.problems[p] = arch_h.value
        self.arch[arch_h.value]["refs"] += 1
        return ctypes.c_void_p(p)

    def dhh_problem_hamiltonian(self, h):
        return h.value + 1

    def dhh_problem_basis(self, h):
        return h.value + 2

    def dhh_problem_destroy(self, h):
        a = self.problems.pop(h.value)
        # the host problem's destructor releases its Hamiltonian and basis through the context
        self._use_ctx(self.arch[a]["ctx"], "problem_destroy")
        self._release_arch(a)
        return 0

    def dhh_last_error(self):
        return b""

    # -- device layer
    def dh_basis_dimension(self, b):
        return 4

    def dh_last_error(self, c):
        return b""

    def dh_malloc(self, c, out, nbytes):
        self._use_ctx(c, "malloc")
        p = self._new()
        self.bufs[p] = c
        self.ctx[c]["refs"] += 1
        out._obj.value = p
        return 0

    def dh_free(self, c, p):
        self._use_ctx(c, "free")
        if self.bufs.pop(p, None) != c:
            self.faults.append(("free", "not a buffer of this context"))
            return capi.DH_E_ARG
        self._release_ctx(c)
        return 0


@pytest.fixture
def fake(monkeypatch):
    def install(counting=True):
        f = FakeNative(counting)
        ns = types.SimpleNamespace(**{k: getattr(f, k) for k in dir(f) if k.startswith("dh")})
        monkeypatch.setattr(capi, "_gpu", ns)
        monkeypatch.setattr(capi, "_host", ns)
        return f
    return install


def _orphan_cycle(f):
    """An Arch with a buffer and a problem that reference it back, reachable from nowhere else."""
    a = capi.Arch(0)
    b = capi.DeviceBuffer(a, 16)
    p = capi.Problem(a, f.new_problem(a.h), "sphere")
    a.cycle = (b, p)


def _collect_orphans():
    gc.collect()
    gc.collect()


def test_collected_cycle_reaches_a_freed_context_without_counting(fake):
    """The fault at exit: weak references are cleared before finalizers run, so the Arch finds
    none of its objects to close, and when the context is freed at dhh_arch_destroy the others'
    finalizers reach it afterwards."""
    f = fake(counting=False)
    _orphan_cycle(f)
    _collect_orphans()
    assert f.log.index("arch_destroy") < max(f.log.index("free"), f.log.index("problem_destroy"))
    assert any(msg == "context used after free" for _, msg in f.faults)


def test_collected_cycle_releases_everything_with_counting(fake):
    f = fake(counting=True)
    _orphan_cycle(f)
    _collect_orphans()
    assert f.faults == []
    assert not f.bufs and not f.problems
    assert all(c["freed"] for c in f.ctx.values())


def test_close_releases_dependents_then_drops_handles(fake):
    f = fake()
    a = capi.Arch(0)
    b = capi.DeviceBuffer(a, 16)
    p = capi.Problem(a, f.new_problem(a.h), "sphere")
    a.close()
    assert f.faults == [] and not f.bufs and not f.problems
    assert all(c["freed"] for c in f.ctx.values())
    assert a.h is None and a.ctx is None
    assert p.h is None and p.ham is None and p.basis is None
    assert b.ptr is None
    # closing twice, and closing the objects again, reach nothing native
    n = len(f.log)
    a.close()
    b.free()
    p.close()
    assert len(f.log) == n and f.faults == []


def test_buffer_freed_after_its_arch(fake):
    """A buffer released after its Arch still frees in its own context, which it keeps alive."""
    f = fake()
    a = capi.Arch(0)
    b = capi.DeviceBuffer(a, 16)
    a._dependents.discard(b)  # as if the Arch could not find it
    a.close()
    assert not any(c["freed"] for c in f.ctx.values())
    b.free()
    assert f.faults == [] and all(c["freed"] for c in f.ctx.values())


def test_exit_hook_closes_open_archs(fake):
    f = fake()
    a = capi.Arch(0)
    capi.DeviceBuffer(a, 16)
    capi._close_live_archs()
    assert a.h is None and f.faults == [] and not f.bufs
    assert all(c["freed"] for c in f.ctx.values())
