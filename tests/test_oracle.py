"""Pins of the CPU oracle (oracle/dh_oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible slip in the restatement (a dropped sign, a wrong
orbital index, a transposed operand, a missing factor 2) fails: reference outputs (golden),
a documented known-answer basis, closed forms, and brute force on tiny systems built here
independently of the oracle's code."""
import itertools
import math

import numpy as np
import pytest

import oracle_lib as O


def coulomb(lz_max):
    return O.round14(O.coulomb_pseudopotentials(lz_max))


# ---------------------------------------------------------------------------- bases
def test_doc_fixture_basis(golden):
    g = golden["fermion_sphere_doc_basis_n4_2s8_lz1"]
    s = O.FermionSphere(g["nbr_fermions"], g["total_lz"], g["lz_max"])
    expect = [sum(1 << i for i, o in enumerate(row) if o) for row in g["occupations"]]
    assert [int(x) for x in s.states()] == expect


def brute_sector(n, total_lz, lz_max):
    """All n-subsets of lz_max+1 orbitals with the sector's index sum, descending."""
    target = (total_lz + n * lz_max) // 2
    out = [sum(1 << i for i in c) for c in itertools.combinations(range(lz_max + 1), n) if sum(c) == target]
    return sorted(out, reverse=True)


@pytest.mark.parametrize("n,total_lz,lz_max", [(4, 2, 8), (3, 1, 6), (5, 0, 11), (6, -4, 12), (2, 0, 4)])
def test_basis_brute_force(n, total_lz, lz_max):
    s = O.FermionSphere(n, total_lz, lz_max)
    assert [int(x) for x in s.states()] == brute_sector(n, total_lz, lz_max)
    lzm = s.state_lzmax()
    for st, m in zip(s.states(), lzm):
        assert int(st).bit_length() - 1 == m


def test_dimensions(golden):
    assert O.lib().or_fs_dimension(8, 21, 0) == golden["fermions_sphere_n8_2s21_lz0"]["dimension"]
    assert O.lib().or_fs_dimension(10, 27, 0) == golden["fermions_sphere_n10_2s27_lz0"]["dimension"]
    assert O.lib().or_fs_dimension(12, 33, 0) == 7764392
    # sum over all Lz sectors = C(2S+1, N)
    tot = sum(O.lib().or_fs_dimension(8, 21, lz) for lz in range(-8 * 21, 8 * 21 + 1, 2))
    assert tot == math.comb(22, 8)


def test_find_state_index_round_trip():
    s = O.FermionSphere(8, 0, 21)
    st, lzm = s.states(), s.state_lzmax()
    assert all(s.find(st[i], lzm[i]) == i for i in range(s.dim))


# ---------------------------------------------------------------------------- operators (brute force)
def jw_apply(state, ops):
    """Apply a product of fermion operators (rightmost first) by Jordan-Wigner with the
    highest orbital first in the normal order: sign = (-1)^(number of occupied orbitals above)."""
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


def test_adadaa_against_jordan_wigner():
    s = O.FermionSphere(4, 2, 8)
    st = s.states()
    index = {int(v): i for i, v in enumerate(st)}
    checked = 0
    for i in range(s.dim):
        for m1, m2, n1, n2 in itertools.product(range(9), repeat=4):
            if m1 + m2 != n1 + n2 or m1 == m2 or n1 == n2:
                continue
            idx, c = s.adadaa(i, m1, m2, n1, n2)
            tgt, sg = jw_apply(int(st[i]), [("c", m1), ("c", m2), ("a", n1), ("a", n2)])
            if tgt is None:
                assert c == 0.0
            else:
                assert idx == index[tgt] and c == sg
                checked += 1
    assert checked > 400


def test_aa_adad_against_jordan_wigner():
    s = O.FermionSphere(4, 2, 8)
    st = s.states()
    index = {int(v): i for i, v in enumerate(st)}
    for i in range(s.dim):
        for n1 in range(9):
            for n2 in range(n1):
                c1 = s.aa(i, n1, n2)
                mid, sg1 = jw_apply(int(st[i]), [("a", n1), ("a", n2)])
                if mid is None:
                    assert c1 == 0.0
                    continue
                assert c1 == sg1
                for m3 in range(9):
                    m4 = n1 + n2 - m3
                    if not (0 <= m4 < m3 <= 8):
                        continue
                    idx, c2 = s.adad(m3, m4)
                    tgt, sg2 = jw_apply(mid, [("c", m3), ("c", m4)])
                    if tgt is None:
                        assert c2 == 0.0
                    else:
                        assert idx == index[tgt] and c2 == sg2


# ---------------------------------------------------------------------------- CG / pseudopotentials
def racah(j1, m1, j2, m2, J):
    """Racah formula in exact rational arithmetic (doubled arguments); returns a float."""
    from fractions import Fraction
    M = m1 + m2
    if abs(m1) > j1 or abs(m2) > j2 or abs(M) > J or J < abs(j1 - j2) or J > j1 + j2:
        return 0.0
    f = math.factorial
    a, b, c, d = (J + j1 - j2) // 2, (J - j1 + j2) // 2, (j1 + j2 - J) // 2, (j1 + j2 + J) // 2 + 1
    pre = Fraction((J + 1) * f(a) * f(b) * f(c), f(d)) * f((J + M) // 2) * f((J - M) // 2) * \
        f((j1 - m1) // 2) * f((j1 + m1) // 2) * f((j2 - m2) // 2) * f((j2 + m2) // 2)
    s = Fraction(0)
    for k in range(0, c + 1):
        args = [k, c - k, (j1 - m1) // 2 - k, (j2 + m2) // 2 - k, (J - j2 + m1) // 2 + k, (J - j1 - m2) // 2 + k]
        if min(args) < 0:
            continue
        den = 1
        for x in args:
            den *= f(x)
        s += Fraction((-1) ** k, den)
    val = s * s * pre
    return float((1 if s >= 0 else -1) * math.sqrt(val))


@pytest.mark.parametrize("j1,j2", [(21, 21), (8, 8), (7, 4), (27, 27)])
def test_clebsch_gordan_vs_racah(j1, j2):
    cg = O.lib().or_cg_create(j1, j2)
    rng = np.random.default_rng(j1 * 100 + j2)
    for _ in range(300):
        m1 = int(rng.integers(0, j1 + 1)) * 2 - j1
        m2 = int(rng.integers(0, j2 + 1)) * 2 - j2
        jmin = max(abs(j1 - j2), abs(m1 + m2))
        J = jmin + 2 * int(rng.integers(0, (j1 + j2 - jmin) // 2 + 1))
        assert abs(O.lib().or_cg_get(cg, m1, m2, J) - racah(j1, m1, j2, m2, J)) < 1e-13
    O.lib().or_cg_destroy(cg)


def test_clebsch_gordan_orthonormal():
    j1 = j2 = 13
    cg = O.lib().or_cg_create(j1, j2)
    for M in range(-26, 27, 2):
        Js = [J for J in range(abs(M), 27, 2)]
        ms = [m1 for m1 in range(-13, 14, 2) if abs(M - m1) <= 13]
        C = np.array([[O.lib().or_cg_get(cg, m1, M - m1, J) for J in Js] for m1 in ms])
        assert np.allclose(C.T @ C, np.eye(len(Js)), atol=1e-13)
    O.lib().or_cg_destroy(cg)


@pytest.mark.parametrize("lz", [8, 21, 27, 33])
def test_pseudopotentials_cg_sum_equals_closed_form(lz):
    a = O.coulomb_pseudopotentials(lz)
    b = O.coulomb_closed_form(lz)
    assert np.max(np.abs(a - b) / np.abs(b)) < 1e-13


def test_pseudopotential_v1_value():
    """V_1 at 2S = 21 from the closed form evaluated in exact integers."""
    S2, m = 21, 1
    L = S2 - m
    v = 2 * math.comb(2 * S2 - 2 * L, S2 - L) * math.comb(2 * S2 + 2 * L + 2, S2 + L + 1) / math.comb(2 * S2 + 2, S2 + 1) ** 2
    v /= math.sqrt(S2 / 2)
    assert abs(O.coulomb_pseudopotentials(21)[1] - v) < 1e-14


# ---------------------------------------------------------------------------- factors and H.v
def test_factor_counts(golden):
    for lz, key in [(21, "fermions_sphere_n8_2s21_lz0"), (27, "fermions_sphere_n10_2s27_lz0")]:
        n_orb = lz + 1
        # "nbr interaction" printed by EvaluateInteractionFactors counts every (m1 > m2, m3)
        printed = sum(min(a + b, lz) - max(a + b - lz, 0) + 1 for a in range(n_orb) for b in range(a))
        assert printed == golden[key]["nbr_interaction_printed"]
        f = O.Factors(lz, coulomb(lz))
        m1, m2, nm3, m3, F = f.tables()
        assert f.nbr_factors == sum(min(a + b, lz) - (a + b) // 2 for a in range(n_orb) for b in range(a))
        assert np.all(2 * m3 > np.repeat(m1 + m2, nm3))


def test_factor_symmetry():
    """F(m1 m2 -> m3 m4) = F(m3 m4 -> m1 m2): the Hamiltonian is symmetric."""
    lz = 15
    f = O.Factors(lz, coulomb(lz))
    m1, m2, nm3, m3, F = f.tables()
    table = {}
    pos = 0
    for a in range(len(m1)):
        for k in range(nm3[a]):
            c = m3[pos]
            table[(m1[a], m2[a], c, m1[a] + m2[a] - c)] = F[pos]
            pos += 1
    for (a, b, c, d), v in table.items():
        if (c, d, a, b) in table:
            assert abs(table[(c, d, a, b)] - v) < 1e-14


def test_element_counts(golden):
    g = golden["fermions_sphere_n8_2s21_lz0"]
    s = O.FermionSphere(8, 0, 21)
    f = O.Factors(21, coulomb(21))
    assert O.fs_count_elements(s, f) == g["full_matrix_elements"]
    diag = 0
    for r in range(s.dim):
        t, _ = O.fs_row_elements(s, f, r)
        diag += int(np.sum(t == r))
    assert diag == g["diagonal_elements"]
    assert (g["full_matrix_elements"] - diag) // 2 + diag == g["stored_hermitian_elements"]


def brute_sphere_term_counts(states, lz_max):
    """(all terms, diagonal terms) of the two-body product by enumeration independent of the oracle's
    operators: every occupied pair n1 > n2 of a state, every created pair m3 > m4 with the same
    orbital sum and both orbitals free once n1 and n2 are emptied"""
    total = diag = 0
    for s in (int(v) for v in states):
        occ = [o for o in range(lz_max + 1) if (s >> o) & 1]
        for i, n1 in enumerate(occ):
            for n2 in occ[:i]:
                t = s ^ (1 << n1) ^ (1 << n2)
                S = n1 + n2
                for m3 in range(S // 2 + 1, min(S, lz_max) + 1):
                    m4 = S - m3
                    if not (t >> m3) & 1 and not (t >> m4) & 1:
                        total += 1
                        diag += (m3, m4) == (n1, n2)
    return total, diag


def test_element_counts_by_enumeration(golden):
    """The reference's printed counts of the N=8, 2S=21 product (full elements, diagonal ones) and of
    the periodic Heisenberg chains (one element per antiparallel bond), from the bases alone: these
    are survey records of the reference programs, pinned here without any product code."""
    g = golden["fermions_sphere_n8_2s21_lz0"]
    total, diag = brute_sphere_term_counts(O.FermionSphere(8, 0, 21).states(), 21)
    assert (total, diag) == (g["full_matrix_elements"], g["diagonal_elements"])
    for L, key in [(24, "heisenberg_periodic_l24_sz0"), (28, "heisenberg_periodic_l28_sz0")]:
        st = O.SpinChain(L, 0).states() & np.uint64((1 << L) - 1)
        rot = ((st >> np.uint64(1)) | ((st & np.uint64(1)) << np.uint64(L - 1)))
        flips = int(np.unpackbits((st ^ rot).astype(">u8").view(np.uint8)).sum())
        assert flips == golden[key]["offdiagonal_elements"]


def dense_from_hv(apply, dim):
    H = np.zeros((dim, dim))
    for i in range(dim):
        e = np.zeros(dim)
        e[i] = 1.0
        H[:, i] = apply(e)
    return H


def second_quantised_dense(n, total_lz, lz_max, pseudo):
    """Dense H built independently: H = sum over (m1>m2 annihilated, m3>m4 created) of F a+_m3 a+_m4 a_m1 a_m2
    with F from the exact Racah CGs (not the oracle's recursion)."""
    states = brute_sector(n, total_lz, lz_max)
    index = {s: i for i, s in enumerate(states)}
    D = len(states)
    H = np.zeros((D, D))
    parity = 0 if lz_max & 1 else 1
    for m1 in range(lz_max + 1):
        for m2 in range(m1):
            for m3 in range(lz_max + 1):
                m4 = m1 + m2 - m3
                if not (0 <= m4 < m3 <= lz_max):
                    continue
                d = [2 * x - lz_max for x in (m1, m2, m3, m4)]
                F = 0.0
                for J in range(abs(d[0] + d[1]), 2 * lz_max + 1, 2):
                    if (J // 2) & 1 != parity:
                        continue
                    F += pseudo[lz_max - J // 2] * racah(lz_max, d[0], lz_max, d[1], J) * racah(lz_max, d[2], lz_max, d[3], J)
                F *= -4.0
                for i, s in enumerate(states):
                    t, sg = jw_apply(s, [("c", m3), ("c", m4), ("a", m1), ("a", m2)])
                    if t is not None:
                        H[index[t], i] += sg * F
    return H


def test_hv_against_second_quantised_brute_force():
    n, tlz, lz = 4, 0, 7
    V = coulomb(lz)
    s = O.FermionSphere(n, tlz, lz)
    f = O.Factors(lz, V)
    H = dense_from_hv(lambda e: O.fs_hv(s, f, e), s.dim)
    Hb = second_quantised_dense(n, tlz, lz, V)
    assert np.allclose(H, H.T, atol=1e-14)
    assert np.max(np.abs(H - Hb)) < 1e-13


def test_hermitian_and_smp_forms_agree():
    s = O.FermionSphere(8, 0, 21)
    f = O.Factors(21, coulomb(21))
    fh = O.Factors(21, coulomb(21), hermitian=True)
    x = O.splitmix(3, s.dim)
    y = O.fs_hv(s, f, x)
    assert np.linalg.norm(O.fs_hv(s, fh, x, hermitian=True) - y) / np.linalg.norm(y) < 1e-13
    assert np.linalg.norm(O.fs_hv(s, f, x, threads=4) - y) / np.linalg.norm(y) < 1e-13
    x2 = O.splitmix(4, s.dim)
    assert abs(x2 @ y - x @ O.fs_hv(s, f, x2)) < 1e-10 * abs(x2 @ y)


def test_row_gather_form_equals_scatter():
    """or_fs_hv_rows (each row gathers its own elements, used to check slices of large sectors)
    equals the reference's scatter loop, whole and on a slice, with a shift."""
    s = O.FermionSphere(8, 0, 21)
    f = O.Factors(21, coulomb(21))
    x = O.splitmix(5, s.dim)
    y = O.fs_hv(s, f, x, shift=0.75)
    assert np.linalg.norm(O.fs_hv_rows(s, f, x, 0, s.dim, shift=0.75) - y) <= 1e-13 * np.linalg.norm(y)
    a, b = 1234, 4321
    assert np.linalg.norm(O.fs_hv_rows(s, f, x, a, b, shift=0.75) - y[a:b]) <= 1e-13 * np.linalg.norm(y[a:b])
    # three vectors in [row][k] layout: each column is the one-vector product
    X = np.stack([x, O.splitmix(6, s.dim), O.splitmix(7, s.dim)], axis=1)
    Y = O.fs_hv_rows(s, f, X.ravel(), a, b, shift=0.75, k=3).reshape(-1, 3)
    for v in range(3):
        yv = O.fs_hv(s, f, X[:, v], shift=0.75)[a:b]
        assert np.linalg.norm(Y[:, v] - yv) <= 1e-13 * np.linalg.norm(yv)


def test_stored_rows_equal_the_product():
    s = O.FermionSphere(8, 0, 21)
    f = O.Factors(21, coulomb(21))
    m = O.StoredRows(s, f, threads=4)
    assert m.nnz == O.fs_count_elements(s, f)
    x = O.splitmix(8, s.dim)
    y = O.fs_hv(s, f, x)
    assert np.linalg.norm(m.hv(x) - y) <= 1e-13 * np.linalg.norm(y)


def test_lanczos_matches_dense_small():
    n, tlz, lz = 5, 1, 11
    s = O.FermionSphere(n, tlz, lz)
    f = O.Factors(lz, coulomb(lz))
    H = dense_from_hv(lambda e: O.fs_hv(s, f, e), s.dim)
    w = np.linalg.eigvalsh(H)
    ev, it = O.fs_lanczos(s, f, 5, threads=2)
    assert it > 0
    assert np.max(np.abs(ev - w[:5])) < 1e-11


@pytest.mark.slow
def test_lanczos_reference_eigenvalues_n8_2s21(golden):
    ref = np.array(golden["fermions_sphere_n8_2s21_lz0"]["eigenvalues"])
    s = O.FermionSphere(8, 0, 21)
    f = O.Factors(21, coulomb(21))
    ev, it = O.fs_lanczos(s, f, 10)
    assert it > 0
    assert np.max(np.abs(ev - ref)) < 1e-10


# ---------------------------------------------------------------------------- spin chain
def test_spin_basis_brute_force():
    for L, sz in [(6, 0), (8, 2), (9, -1)]:
        s = O.SpinChain(L, sz)
        nup = (L + sz) // 2
        states = []
        for ups in itertools.combinations(range(L), nup):
            w = 0xFFFFFFFF
            for i in range(L):
                if i not in ups:
                    w &= ~(1 << i)
            states.append(w)
        key = lambda w: int(format(w & ((1 << L) - 1), "0%db" % L)[::-1], 2)  # site 0 most significant
        states.sort(key=key, reverse=True)
        assert [int(x) for x in s.states()] == states
        assert all(s.find(int(x)) == i for i, x in enumerate(s.states()))


def heisenberg_dense_independent(L, sz, periodic=True):
    nup = (L + sz) // 2
    states = sorted([sum(1 << i for i in c) for c in itertools.combinations(range(L), nup)])
    index = {s: i for i, s in enumerate(states)}
    D = len(states)
    H = np.zeros((D, D))
    bonds = [(i, i + 1) for i in range(L - 1)] + ([(L - 1, 0)] if periodic else [])
    for i, s in enumerate(states):
        for a, b in bonds:
            if ((s >> a) ^ (s >> b)) & 1:
                H[i, i] -= 0.25
                H[index[s ^ (1 << a) ^ (1 << b)], i] += 0.5
            else:
                H[i, i] += 0.25
    return np.linalg.eigvalsh(H)


@pytest.mark.parametrize("L,periodic", [(10, True), (9, False), (12, True)])
def test_spin_spectrum_vs_independent_dense(L, periodic):
    s = O.SpinChain(L, 0 if L % 2 == 0 else 1)
    H = dense_from_hv(lambda e: s.hv(e, periodic=periodic), s.dim)
    assert np.allclose(H, H.T)
    w = np.linalg.eigvalsh(H)
    assert np.max(np.abs(w - heisenberg_dense_independent(L, 0 if L % 2 == 0 else 1, periodic))) < 1e-12
    if L == 10 and periodic:
        assert abs(w[0] - (-4.515446354492)) < 1e-9  # E0 of the 10-site periodic chain


@pytest.mark.slow
def test_spin_lanczos_l16_vs_dense_free_values():
    s = O.SpinChain(16, 0)
    ev, it = s.lanczos(2)
    assert it > 0
    # E0 of the 16-site periodic Heisenberg chain
    assert abs(ev[0] - (-7.142296360817)) < 1e-9


def test_tridiagonal_ql_matches_numpy():
    rng = np.random.default_rng(5)
    n = 40
    d = rng.standard_normal(n)
    u = rng.standard_normal(n)
    T = np.diag(d) + np.diag(u[:-1], 1) + np.diag(u[:-1], -1)
    dd, uu = d.copy(), u.copy()
    uu[-1] = 0.0
    O.lib().or_tridiag_eigenvalues(n, O.dptr(dd), O.dptr(uu), 50)
    assert np.max(np.abs(np.sort(dd) - np.linalg.eigvalsh(T))) < 1e-12
