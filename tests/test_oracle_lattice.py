"""Pins of the oracle's momentum-space (Chern band) path: basis against brute force, the
projected two-body Hamiltonian against an independent construction from the real-space
nearest-neighbour interaction, Hermiticity and gauge invariance of the spectrum."""
import itertools

import numpy as np
import pytest

import oracle_lib as O


def brute_momentum_basis(n, nx, ny, kx, ky):
    out = []
    for c in itertools.combinations(range(nx * ny), n):
        if sum(o // ny for o in c) % nx == kx and sum(o % ny for o in c) % ny == ky:
            out.append(sum(1 << o for o in c))
    return sorted(out, reverse=True)


@pytest.mark.parametrize("n,nx,ny,kx,ky", [(3, 3, 2, 0, 0), (3, 3, 2, 1, 1), (4, 4, 3, 2, 0), (2, 2, 2, 1, 0)])
def test_momentum_basis_brute_force(n, nx, ny, kx, ky):
    s = O.MomentumFermions(n, nx, ny, kx, ky)
    assert [int(x) for x in s.states()] == brute_momentum_basis(n, nx, ny, kx, ky)


def test_momentum_dimension_cfg():
    # 6x5 lattice, N = 10: every momentum sector has about C(30,10)/30 states
    total = 0
    dims = []
    for kx in range(6):
        for ky in range(5):
            d = O.lib().or_fs_dim(O.MomentumFermions(10, 6, 5, kx, ky).h) if (kx, ky) in [(0, 0), (1, 2)] else None
            if d:
                dims.append(d)
    assert all(1001400 <= d <= 1001603 for d in dims)


def test_lower_band_is_lowest_eigenvector():
    nx, ny = 6, 5
    u, e = O.checkerboard_lower_band(nx, ny, 1.0, 0.5, -0.5)
    for idx in range(nx * ny):
        Kx, Ky = 2 * np.pi * (idx // ny) / nx, 2 * np.pi * (idx % ny) / ny
        B1 = 4 * (np.cos(Kx / 2) * np.cos(Ky / 2) * np.cos(np.pi / 4) + 1j * np.sin(Kx / 2) * np.sin(Ky / 2) * np.sin(np.pi / 4))
        d1 = -2.0 * np.cos(Kx) * np.cos(Ky)
        d3 = 1.0 * (np.cos(Kx) - np.cos(Ky))
        H = np.array([[d1 + d3, B1], [np.conj(B1), d1 - d3]])
        w, V = np.linalg.eigh(H)
        v = np.conj(u[idx])  # OneBodyBasis stores conj(v)
        assert abs(np.linalg.norm(v) - 1) < 1e-14
        assert np.linalg.norm(H @ v - w[0] * v) < 1e-13
        assert abs(e[idx, 0] - w[0]) < 1e-13 and abs(e[idx, 1] - w[1]) < 1e-13


def jw(state, ops):
    sign = 1
    for kind, m in reversed(ops):
        above = bin(state >> (m + 1)).count("1")
        occ = (state >> m) & 1
        if kind == "a":
            if not occ:
                return None, 0
            state &= ~(1 << m)
        else:
            if occ:
                return None, 0
            state |= 1 << m
        sign *= -1 if above & 1 else 1
    return state, sign


def projected_nn_hamiltonian(n, nx, ny, kx, ky, v):
    """Independent construction: U sum_<A,B> n_A n_B on the checkerboard lattice (B at A + (+-1/2, +-1/2)),
    Fourier with site positions, projected on the band whose Bloch vectors are v[k] (U = 1):
    H = (1/Nc) sum_{k1+k2=k3+k4} f(k4-k2) conj(vA(k3)) conj(vB(k4)) vB(k2) vA(k1) g+_k3 g+_k4 g_k2 g_k1,
    f(q) = sum_delta exp(i q.delta) = 4 cos(qx/2) cos(qy/2)."""
    Nc = nx * ny
    states = brute_momentum_basis(n, nx, ny, kx, ky)
    index = {s: i for i, s in enumerate(states)}
    D = len(states)
    H = np.zeros((D, D), complex)
    K = [(2 * np.pi * (o // ny) / nx, 2 * np.pi * (o % ny) / ny) for o in range(Nc)]
    for k1 in range(Nc):
        for k2 in range(Nc):
            for k3 in range(Nc):
                k4x = (k1 // ny + k2 // ny - k3 // ny) % nx
                k4y = (k1 % ny + k2 % ny - k3 % ny) % ny
                k4 = k4x * ny + k4y
                # q = k4 - k2 taken as the lattice momentum difference (f is 2 pi periodic only up to the
                # sign of the half-phase, so use the actual difference of the representatives)
                qx = K[k4][0] - K[k2][0]
                qy = K[k4][1] - K[k2][1]
                f = 2.0 * (np.cos(0.5 * (qx - qy)) + np.cos(0.5 * (qx + qy)))
                amp = f * np.conj(v[k3][0]) * np.conj(v[k4][1]) * v[k2][1] * v[k1][0] / Nc
                if abs(amp) < 1e-15:
                    continue
                for i, s in enumerate(states):
                    t, sg = jw(s, [("c", k3), ("c", k4), ("a", k2), ("a", k1)])
                    if t is not None:
                        H[index[t], i] += sg * amp
    return H


@pytest.mark.parametrize("n,nx,ny,kx,ky", [(3, 3, 2, 0, 0), (3, 3, 2, 2, 1), (2, 3, 3, 1, 2)])
def test_projected_hamiltonian_spectrum_matches_independent_construction(n, nx, ny, kx, ky):
    u, _ = O.checkerboard_lower_band(nx, ny, 1.0, 0.5, -0.5)
    space = O.MomentumFermions(n, nx, ny, kx, ky)
    fac = O.LatticeFactors(nx, ny, u)
    D = space.dim
    H = np.zeros((D, D), complex)
    for c in range(D):
        e = np.zeros(D, complex)
        e[c] = 1.0
        H[:, c] = O.fm_hv(space, fac, e)
    assert np.max(np.abs(H - H.conj().T)) < 1e-13
    Hh = np.zeros((D, D), complex)
    for c in range(D):
        e = np.zeros(D, complex)
        e[c] = 1.0
        Hh[:, c] = O.fm_hv(space, fac, e, hermitian=True)
    assert np.max(np.abs(H - Hh)) < 1e-13
    Hb = projected_nn_hamiltonian(n, nx, ny, kx, ky, np.conj(u))
    assert np.max(np.abs(Hb - Hb.conj().T)) < 1e-12
    w = np.linalg.eigvalsh(H)
    wb = np.linalg.eigvalsh(Hb)
    assert np.max(np.abs(w - wb)) < 1e-12, (w[:4], wb[:4])


def test_row_gather_form_equals_scatter():
    """or_fm_hv_rows (rows gather the conjugates of the elements their own row produces) equals the
    reference's scatter loop on a complex vector, whole and on a slice."""
    u, _ = O.checkerboard_lower_band(4, 3, 1.0, 0.5, -0.5)
    space = O.MomentumFermions(4, 4, 3, 1, 2)
    fac = O.LatticeFactors(4, 3, u)
    rng = np.random.default_rng(8)
    x = rng.standard_normal(space.dim) + 1j * rng.standard_normal(space.dim)
    y = O.fm_hv(space, fac, x)
    assert np.linalg.norm(O.fm_hv_rows(space, fac, x, 0, space.dim) - y) <= 1e-13 * np.linalg.norm(y)
    a, b = 17, space.dim - 9
    assert np.linalg.norm(O.fm_hv_rows(space, fac, x, a, b) - y[a:b]) <= 1e-13 * np.linalg.norm(y[a:b])
    X = np.stack([x, rng.standard_normal(space.dim) + 1j * rng.standard_normal(space.dim)], axis=1)
    Y = O.fm_hv_rows(space, fac, X.ravel(), a, b, k=2).reshape(-1, 2)
    for v in range(2):
        yv = O.fm_hv(space, fac, X[:, v])[a:b]
        assert np.linalg.norm(Y[:, v] - yv) <= 1e-13 * np.linalg.norm(yv)


def test_spectrum_is_gauge_invariant():
    nx, ny, n = 3, 2, 3
    u, _ = O.checkerboard_lower_band(nx, ny, 1.0, 0.5, -0.5)
    rng = np.random.default_rng(0)
    phases = np.exp(1j * rng.uniform(0, 2 * np.pi, nx * ny))
    space = O.MomentumFermions(n, nx, ny, 0, 0)
    spectra = []
    for uu in (u, u * phases[:, None]):
        fac = O.LatticeFactors(nx, ny, uu)
        D = space.dim
        H = np.zeros((D, D), complex)
        for c in range(D):
            e = np.zeros(D, complex)
            e[c] = 1.0
            H[:, c] = O.fm_hv(space, fac, e)
        spectra.append(np.linalg.eigvalsh(H))
    assert np.max(np.abs(spectra[0] - spectra[1])) < 1e-13
