"""Adjoint kernel: bilinears <psi|H_m|phi> per interaction coefficient and the VJP of H.x,
wrapped in torch.autograd.Function (test harness only) and checked against the oracle and
against central finite differences of oracle Lanczos energies."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def channel_matrix_vector(space, lz_max, m, x):
    """H_m x from the oracle: factors of the pseudopotential unit vector e_m."""
    e = np.zeros(lz_max + 1)
    e[m] = 1.0
    return O.fs_hv(space, O.Factors(lz_max, e), x)


@pytest.mark.parametrize("n,total_lz,lz,same", [(6, 0, 15, False), (6, 0, 15, True), (4, 2, 33, False),
                                                 (5, 0, 20, True), (2, 0, 9, False)])
def test_bilinears_match_oracle(arch, n, total_lz, lz, same):
    """Every channel's <psi|H_m|phi> against the oracle's product with the unit pseudopotential e_m:
    psi and phi distinct and equal (the expectation-value form), 8 to 17 channels (two groups of
    channels at 2S = 33), and N = 2 (the (N-2)-particle state is empty)."""
    from diagham_amd import capi
    V = capi.coulomb_pseudopotentials(lz, True)
    p = capi.Problem.sphere(arch, n, total_lz, lz, V)
    space = O.FermionSphere(n, total_lz, lz)
    psi = O.splitmix(31, p.dim)
    phi = psi if same else O.splitmix(32, p.dim)
    dpsi = capi.DeviceBuffer.from_host(arch, psi)
    dphi = dpsi if same else capi.DeviceBuffer.from_host(arch, phi)
    dh = capi.DeviceBuffer(arch, p.dim)
    g = p.adjoint_device(dpsi.ptr, dphi.ptr, dh.ptr)
    mom = p.channels()
    assert list(mom) == list(range(1, lz + 1, 2))
    for c, m in enumerate(mom):
        ref = psi @ channel_matrix_vector(space, lz, m, phi)
        assert abs(g[c] - ref) <= 1e-12 * max(1.0, abs(ref)), (m, g[c], ref)
    # the same pass returns H psi
    hpsi = dh.download()
    ref = O.fs_hv(space, O.Factors(lz, V), psi)
    assert np.linalg.norm(hpsi - ref) / np.linalg.norm(ref) < 1e-12
    # linearity: sum_m V_m <psi|H_m|phi> = <psi|H|phi>
    assert abs(np.dot(V[mom], g) - psi @ O.fs_hv(space, O.Factors(lz, V), phi)) < 1e-10
    p.close()


def _torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("torch has no GPU")
    return torch


def make_functions(torch, arch, capi, n, total_lz, lz):
    """torch.autograd.Function wrappers over the C ABI (test harness only)."""

    class HV(torch.autograd.Function):
        """y = H(V) x;  backward: x_bar = H y_bar, V_bar[m] = <y_bar|H_m|x>."""

        @staticmethod
        def forward(ctx, V, x):
            p = capi.Problem.sphere(arch, n, total_lz, lz, V.detach().cpu().numpy())
            y = torch.empty_like(x)
            torch.cuda.synchronize()
            p.hv_device(x.data_ptr(), y.data_ptr(), 1)
            arch.sync()
            ctx.problem = p
            ctx.save_for_backward(x)
            return y

        @staticmethod
        def backward(ctx, ybar):
            (x,) = ctx.saved_tensors
            p = ctx.problem
            ybar = ybar.contiguous()
            xbar = torch.empty_like(x)
            torch.cuda.synchronize()
            g = p.adjoint_device(ybar.data_ptr(), x.data_ptr(), xbar.data_ptr())
            arch.sync()
            Vbar = torch.zeros(lz + 1, dtype=torch.float64)
            Vbar[torch.tensor(p.channels(), dtype=torch.long)] = torch.tensor(g)
            p.close()
            return Vbar, xbar

    class Energies(torch.autograd.Function):
        """Lowest eigenvalues of H(V) (GPU Lanczos); backward by Hellmann-Feynman:
        dE_i/dV_m = <psi_i|H_m|psi_i> from the adjoint kernel."""

        @staticmethod
        def forward(ctx, V, nbr):
            p = capi.Problem.sphere(arch, n, total_lz, lz, V.detach().cpu().numpy())
            vecs = torch.empty((nbr, p.dim), dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            r = p.lanczos(nbr, block_size=1, max_iter=800, residual=1e-11, d_vectors=vecs.data_ptr(),
                          nbr_vectors=nbr)
            assert r["converged"], "Lanczos did not converge within 800 products"
            G = torch.zeros((nbr, lz + 1), dtype=torch.float64)
            idx = torch.tensor(p.channels(), dtype=torch.long)
            for i in range(nbr):
                g = p.adjoint_device(vecs[i].data_ptr(), vecs[i].data_ptr())
                G[i, idx] = torch.tensor(g)
            ctx.save_for_backward(G)
            p.close()
            return torch.tensor(r["eigenvalues"])

        @staticmethod
        def backward(ctx, gE):
            (G,) = ctx.saved_tensors
            return gE @ G, None

    return HV, Energies


def test_vjp_of_product(arch):
    torch = _torch()
    from diagham_amd import capi
    n, lz = 5, 13
    arch.check(capi.gpu().dh_set_stream(arch.ctx, torch.cuda.current_stream().cuda_stream), "dh_set_stream")
    HV, _ = make_functions(torch, arch, capi, n, 1, lz)
    V = torch.tensor(capi.coulomb_pseudopotentials(lz, True), dtype=torch.float64, requires_grad=True)
    dim = capi.fermion_sphere_dimension(n, 1, lz)
    x = torch.tensor(O.splitmix(5, dim), dtype=torch.float64, device="cuda", requires_grad=True)
    w = torch.tensor(O.splitmix(6, dim), dtype=torch.float64, device="cuda")
    loss = (HV.apply(V, x) * w).sum()
    loss.backward()
    # H is linear in V and x: central differences are exact up to rounding
    space = O.FermionSphere(n, 1, lz)
    xn, wn = x.detach().cpu().numpy(), w.cpu().numpy()
    for m in range(1, lz + 1, 2):
        ref = wn @ channel_matrix_vector(space, lz, m, xn)
        assert abs(V.grad[m].item() - ref) <= 1e-11 * max(1.0, abs(ref))
    for m in range(0, lz + 1, 2):
        assert V.grad[m].item() == 0.0  # even m never enter a fermionic interaction
    Hw = O.fs_hv(space, O.Factors(lz, V.detach().numpy()), wn)
    assert np.linalg.norm(x.grad.cpu().numpy() - Hw) / np.linalg.norm(Hw) < 1e-12


def test_energy_gradients_match_finite_differences(arch):
    """dE0/dV_m and d(E1-E0)/dV_m for N=10, 2S=27, Lz=0 against central differences
    (h = 1e-5) of oracle Lanczos energies (tests/golden/fd_n10_2s27.json), 1e-6 relative."""
    torch = _torch()
    from diagham_amd import capi
    path = os.path.join(HERE, "golden", "fd_n10_2s27.json")
    if not os.path.exists(path):
        pytest.skip("finite-difference fixture not generated")
    fd = json.load(open(path))
    arch.check(capi.gpu().dh_set_stream(arch.ctx, torch.cuda.current_stream().cuda_stream), "dh_set_stream")
    _, Energies = make_functions(torch, arch, capi, fd["nbr_fermions"], fd["total_lz"], fd["lz_max"])
    V = torch.tensor(fd["pseudopotentials"], dtype=torch.float64, requires_grad=True)
    E = Energies.apply(V, 2)
    assert abs(E[0].item() - fd["E0"]) < 1e-10 and abs(E[1].item() - fd["E1"]) < 1e-10
    (g0,) = torch.autograd.grad(E[0], V, retain_graph=True)
    (gg,) = torch.autograd.grad(E[1] - E[0], V)
    for i, m in enumerate(fd["free_m"]):
        a, b = g0[m].item(), fd["dE0_dV"][i]
        assert abs(a - b) <= 1e-6 * abs(b), ("dE0", m, a, b)
        a, b = gg[m].item(), fd["dGap_dV"][i]
        assert abs(a - b) <= 1e-6 * max(abs(b), 1e-3), ("dGap", m, a, b)


def test_l0_gap_gradients_match_finite_differences(arch):
    """The differentiation workload as specified: the two lowest L = 0 states of N=10, 2S=27, Lz=0.
    They are the lowest states of H + c L^2 (the --l2-factor of FQHESphereFermionsTwoBodyGeneric.cc:170,
    every L > 0 state moved up by c L (L + 1)); c = 0.1 lifts the L = 2 and 3 states that lie below
    the second L = 0 state (0.31 above the first) over it, and keeps the spectrum narrow enough for
    the Ritz vectors to converge quickly.  The eigenvectors give dE/dV_m = <psi|H_m|psi> through
    the adjoint of H alone.  Energies against the oracle's L = 0 levels (1e-10), gradients against
    central differences of them (tests/golden/fd_n10_2s27_l0.json, 1e-6 relative)."""
    torch = _torch()
    from diagham_amd import capi
    path = os.path.join(HERE, "golden", "fd_n10_2s27_l0.json")
    if not os.path.exists(path):
        pytest.skip("finite-difference fixture of the L = 0 levels not generated")
    fd = json.load(open(path))
    n, tlz, lz = fd["nbr_fermions"], fd["total_lz"], fd["lz_max"]
    V = np.array(fd["pseudopotentials"])
    shifted = capi.Problem.sphere_l2(arch, n, tlz, lz, 0.1, pseudo=V)
    vecs = torch.empty((2, shifted.dim), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    # the gradients are bilinears of the Ritz vectors, whose error is first order in the residual
    # (the eigenvalues' error is second order): the residual test converges the vectors as well
    r = shifted.lanczos(2, block_size=1, max_iter=3000, residual=1e-11, d_vectors=vecs.data_ptr(),
                        nbr_vectors=2)
    assert r["converged"]
    E = r["eigenvalues"]
    assert abs(E[0] - fd["E0"]) < 1e-10 and abs(E[1] - fd["E1"]) < 1e-10, (E, fd["E0"], fd["E1"])
    L2 = capi.Problem.sphere_l2(arch, n, tlz, lz, 1.0)
    y = torch.empty_like(vecs[0])
    for i in range(2):
        torch.cuda.synchronize()
        L2.hv_device(vecs[i].data_ptr(), y.data_ptr(), 1)
        arch.sync()
        assert abs(torch.dot(vecs[i], y).item()) < 1e-8  # L = 0
    p = capi.Problem.sphere(arch, n, tlz, lz, V)
    mom = list(p.channels())
    g = [p.adjoint_device(vecs[i].data_ptr(), vecs[i].data_ptr()) for i in range(2)]
    arch.sync()
    # 1e-6 relative, plus the finite differences' own noise: each energy of the fixture is known to
    # its Lanczos stopping rule, 1e-14 relative, and a central difference divides that by 2h (the
    # gap's derivative takes four energies); the gap's derivatives are small differences of ~10
    noise = 1e-14 * max(abs(fd["E0"]), abs(fd["E1"])) / (2 * fd["h"])
    for j, m in enumerate(fd["free_m"]):
        c = mom.index(m)
        a, b = g[0][c], fd["dE0_dV"][j]
        assert abs(a - b) <= 1e-6 * abs(b) + 2 * noise, ("dE0", m, a, b)
        a, b = g[1][c] - g[0][c], fd["dGap_dV"][j]
        assert abs(a - b) <= 1e-6 * abs(b) + 4 * noise, ("dGap", m, a, b)
    for q in (shifted, L2, p):
        q.close()
