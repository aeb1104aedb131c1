"""CPU check of the factorised bilinear that the adjoint kernel evaluates: with the channel
interaction F_c(a -> b) = -4 u_c[a] u_c[b] (u_c[p] the Clebsch-Gordan coefficient of pair p in
channel c), <psi|H_c|phi> = 4 sum_t A_c(psi)[t] A_c(phi)[t] over the (N-2)-particle states t, with
A_c(v)[t] = sum over the pairs p free in t whose orbital sum completes t to the sector of
u_c[p] (-1)^(popc(t >> p_lo) + popc(t >> p_hi)) v[t | p].  Both sides from the oracle: the
right-hand side by brute-force enumeration, the left-hand side by the oracle's product with the
unit pseudopotential e_m."""
import itertools

import numpy as np
import pytest

import oracle_lib as O


def popc(x):
    return bin(x).count("1")


@pytest.mark.parametrize("n,total_lz,lz", [(5, 1, 11), (4, 0, 9), (3, 2, 8)])
def test_channel_bilinear_factorises_over_pair_holes(n, total_lz, lz):
    space = O.FermionSphere(n, total_lz, lz)
    states = [int(s) for s in space.states()]
    index = {s: i for i, s in enumerate(states)}
    orbital_sum = sum(j for j in range(lz + 1) if (states[0] >> j) & 1)
    cg = O.lib().or_cg_create(lz, lz)
    rng = np.random.default_rng(n * 100 + lz)
    psi, phi = rng.standard_normal(space.dim), rng.standard_normal(space.dim)
    try:
        for m in range(1, lz + 1, 2):
            e = np.zeros(lz + 1)
            e[m] = 1.0
            ref = psi @ O.fs_hv(space, O.Factors(lz, e), phi)
            g = 0.0
            for holes in itertools.combinations(range(lz + 1), n - 2):
                t = sum(1 << b for b in holes)
                s = orbital_sum - sum(holes)
                a = b = 0.0
                for m3 in range(s // 2 + 1, min(s, lz) + 1):
                    m4 = s - m3
                    if m4 < 0 or m4 >= m3 or (t >> m3) & 1 or (t >> m4) & 1:
                        continue
                    u = O.lib().or_cg_get(cg, 2 * m3 - lz, 2 * m4 - lz, 2 * (lz - m))
                    sign = -1.0 if (popc(t >> m4) + popc(t >> m3)) & 1 else 1.0
                    target = index[t | (1 << m3) | (1 << m4)]
                    a += u * sign * psi[target]
                    b += u * sign * phi[target]
                g += a * b
            assert abs(4.0 * g - ref) <= 1e-12 * max(1.0, abs(ref)), (m, 4.0 * g, ref)
    finally:
        O.lib().or_cg_destroy(cg)
