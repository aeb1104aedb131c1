"""The oracle's n-body operators (ProdA / ProdAd) and n-body product, pinned by brute-force
Jordan-Wigner second quantisation (no GPU)."""
import itertools

import numpy as np

import oracle_lib as O
from test_oracle import brute_sector, dense_from_hv, jw_apply


def test_prod_a_prod_ad_against_jordan_wigner():
    """ProdA annihilates the rightmost orbital first and counts the occupied orbitals at and above
    each one (its own included), so a product of n annihilations carries (-1)^n with respect to
    the normal-ordered Jordan-Wigner sign; ProdAd matches Jordan-Wigner exactly."""
    n, tlz, lz = 5, 1, 8
    s = O.FermionSphere(n, tlz, lz)
    st = s.states()
    index = {int(v): i for i, v in enumerate(st)}
    triples = [c for c in itertools.combinations(range(lz, -1, -1), 3)]  # descending tuples
    checked = 0
    for i in range(s.dim):
        for nt in triples:
            c = s.prod_a(i, nt)
            mid, sg = jw_apply(int(st[i]), [("a", x) for x in nt])
            if mid is None:
                assert c == 0.0
                continue
            assert c == -sg
            for mt in triples:
                if sum(mt) != sum(nt):
                    continue
                idx, c2 = s.prod_ad(mt)
                tgt, sg2 = jw_apply(mid, [("c", x) for x in mt])
                if tgt is None:
                    assert c2 == 0.0 and idx == s.dim
                else:
                    assert idx == index[tgt] and c2 == sg2
                    checked += 1
    assert checked > 300


def test_tuples_follow_reference_order():
    counts, flat = O.nbody_tuples(6, 3)
    assert counts.sum() == 20 and flat.size == 60
    t = flat.reshape(-1, 3)
    assert all(a > b > c for a, b, c in t)
    sums = t.sum(axis=1)
    assert np.all(np.diff(sums) >= 0)  # grouped by sum
    for S in np.unique(sums):
        grp = [tuple(r) for r in t[sums == S]]
        assert grp == sorted(grp, reverse=True)  # lexicographically descending inside a sum


def nbody_dense(n, total_lz, lz_max, nbody, counts, tuples, factors):
    """Dense H = sum_S sum_{i,l} F_S[i][l] (-1)^nbody a+_{M_l} a_{N_i} by Jordan-Wigner."""
    states = brute_sector(n, total_lz, lz_max)
    index = {s: i for i, s in enumerate(states)}
    H = np.zeros((len(states), len(states)))
    t = tuples.reshape(-1, nbody)
    pos, fpos = 0, 0
    for c in counts:
        F = factors[fpos:fpos + c * c].reshape(c, c)
        for i in range(c):
            for l in range(c):
                ops = [("c", x) for x in t[pos + l]] + [("a", x) for x in t[pos + i]]
                for col, s in enumerate(states):
                    tgt, sg = jw_apply(s, ops)
                    if tgt is not None:
                        H[index[tgt], col] += (-1) ** nbody * sg * F[i, l]
        pos += c
        fpos += c * c
    return H


def test_three_body_product_against_brute_force():
    n, tlz, lz = 5, 1, 8
    counts, tuples = O.nbody_tuples(lz + 1, 3)
    F = O.random_nbody_factors(counts, 4)
    s = O.FermionSphere(n, tlz, lz)
    H = dense_from_hv(lambda e: O.nbody_hv(s, 3, counts, tuples, F, e), s.dim)
    Hb = nbody_dense(n, tlz, lz, 3, counts, tuples, F)
    assert np.max(np.abs(H - Hb)) < 1e-12
    assert np.max(np.abs(H - H.T)) < 1e-12  # symmetric factors give a symmetric operator


def test_two_body_limit_matches_pair_product():
    """nbody = 2 with the pair factors of the Coulomb problem reproduces the two-body product
    (the per-pair factor layout regrouped per sum; ProdA's two signs cancel)."""
    lz = 9
    V = O.round14(O.coulomb_pseudopotentials(lz))
    fac = O.Factors(lz, V)
    m1, m2, nm3, m3, Fp = fac.tables()
    counts, tuples = O.nbody_tuples(lz + 1, 2)
    t = tuples.reshape(-1, 2)
    # F_S[i][l] for annihilated pair i = (n1 > n2), created pair l = (m3 > m4)
    table = {}
    pos = 0
    for a in range(len(m1)):
        for j in range(nm3[a]):
            table[(int(m1[a]), int(m2[a]), int(m3[pos + j]))] = Fp[pos + j]
        pos += nm3[a]
    blocks, p = [], 0
    for c in counts:
        Fs = np.zeros((c, c))
        for i in range(c):
            for l in range(c):
                Fs[i, l] = table[(int(t[p + i][0]), int(t[p + i][1]), int(t[p + l][0]))]
        blocks.append(Fs.ravel())
        p += c
    F = np.concatenate(blocks)
    s = O.FermionSphere(4, 0, lz)
    x = O.splitmix(5, s.dim)
    y2 = O.fs_hv(s, fac, x)
    yn = O.nbody_hv(s, 2, counts, tuples, F, x)
    assert np.max(np.abs(y2 - yn)) < 1e-13 * np.max(np.abs(y2))
