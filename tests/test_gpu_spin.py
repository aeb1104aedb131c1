"""GPU parity of the spin-1/2 Heisenberg chain path against the oracle and reference outputs."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L,sz", [(10, 0), (12, 2), (16, 0), (17, -1)])
def test_basis_bit_exact(arch, L, sz):
    from diagham_amd import capi
    p = capi.Problem.spin_chain(arch, L, sz)
    ref = O.SpinChain(L, sz)
    assert p.dim == ref.dim
    assert np.array_equal(p.states(), ref.states())
    for mode in (0, 1):
        assert np.array_equal(p.find_state_index(ref.states(), mode), np.arange(p.dim))
    p.close()


@pytest.mark.parametrize("L,periodic", [(16, True), (16, False), (20, True)])
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("k", [1, 4])
def test_hv_parity(arch, L, periodic, mode, k):
    from diagham_amd import capi
    rng = np.random.default_rng(L + k)
    J = rng.uniform(0.5, 1.5, L)
    Jz = rng.uniform(0.5, 1.5, L)
    p = capi.Problem.spin_chain(arch, L, 0, J, Jz, periodic)
    p.set_launch(1, 256, mode, 1)
    ref = O.SpinChain(L, 0)
    X = np.stack([O.splitmix(200 + v, p.dim) for v in range(k)], axis=1)
    Y = p.hv(X.reshape(-1), k).reshape(p.dim, k)
    for v in range(k):
        x = np.ascontiguousarray(X[:, v])
        y_ref = np.zeros_like(x)
        O.lib().or_spin_hv(ref.h, O.dptr(J), O.dptr(Jz), int(periodic), O.dptr(x), O.dptr(y_ref))
        err = np.linalg.norm(Y[:, v] - y_ref) / np.linalg.norm(y_ref)
        assert err < 1e-12, (L, periodic, mode, k, v, err)
    p.close()


@pytest.mark.parametrize("L,sz,periodic,k", [(16, 0, True, 1), (17, -1, True, 2), (18, 2, False, 1),
                                             (20, 0, True, 8), (13, 3, True, 4)])
def test_hv_parity_uniform_couplings(arch, L, sz, periodic, k):
    """Equal J and Jz on every bond (the kernel's one-coefficient branch), odd and even L, Sz != 0."""
    from diagham_amd import capi
    J = np.full(L, 0.7)
    Jz = np.full(L, 1.3)
    p = capi.Problem.spin_chain(arch, L, sz, J, Jz, periodic)
    p.set_launch(1, 256, 1, 1)
    ref = O.SpinChain(L, sz)
    X = np.stack([O.splitmix(300 + v, p.dim) for v in range(k)], axis=1)
    Y = p.hv(X.reshape(-1), k).reshape(p.dim, k)
    for v in range(k):
        x = np.ascontiguousarray(X[:, v])
        y_ref = np.zeros_like(x)
        O.lib().or_spin_hv(ref.h, O.dptr(J), O.dptr(Jz), int(periodic), O.dptr(x), O.dptr(y_ref))
        err = np.linalg.norm(Y[:, v] - y_ref) / np.linalg.norm(y_ref)
        assert err < 1e-12, (L, sz, periodic, k, v, err)
    p.close()


def test_row_targets_bit_exact(arch):
    from diagham_amd import capi
    L = 18
    p = capi.Problem.spin_chain(arch, L, 0)
    ref = O.SpinChain(L, 0)
    states = ref.states()
    mask = (1 << L) - 1
    for r in range(0, p.dim, 911):
        t, c = p.row_elements(r)
        s = int(states[r]) & mask
        expect = [r]
        for i in range(L):
            j = (i + 1) % L
            if ((s >> i) ^ (s >> j)) & 1:
                flipped = int(states[r]) ^ (1 << i) ^ (1 << j)
                expect.append(ref.find(flipped))
        assert sorted(t.tolist()) == sorted(expect)
    p.close()


def test_element_count_l24(arch, golden):
    from diagham_amd import capi
    p = capi.Problem.spin_chain(arch, 24, 0)
    g = golden["heisenberg_periodic_l24_sz0"]
    assert p.dim == g["dimension"]
    assert p.count_elements() == g["offdiagonal_elements"] + g["dimension"]
    p.close()


def test_lanczos_l24_reference_eigenvalues(arch, golden):
    from diagham_amd import capi
    p = capi.Problem.spin_chain(arch, 24, 0)
    ref = np.array(golden["heisenberg_periodic_l24_sz0"]["eigenvalues"])
    r = p.lanczos(4, block_size=1)
    assert r["converged"]
    np.testing.assert_allclose(r["eigenvalues"], ref, rtol=0, atol=1e-10)
    p.close()
