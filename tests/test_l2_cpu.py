"""L^2 of fermions on the sphere as a two-body term (host factor tables, no GPU): against L^2
built from Jordan-Wigner ladder operators, and against its spectrum L(L+1) with the multiplicities
that the sector dimensions fix."""
import numpy as np
import pytest

from test_oracle import brute_sector, jw_apply


def ladder_l2(n, total_lz, lz_max):
    """Dense L^2 = L^- L^+ + Lz^2 + Lz on the sector, with L^+ = sum_p c_p a+_{p+1} a_p as a matrix
    between the sector and the next one (Jordan-Wigner)."""
    states = brute_sector(n, total_lz, lz_max)
    up = brute_sector(n, total_lz + 2, lz_max) if total_lz + 2 <= n * (lz_max - n + 1) else []
    idx_up = {s: i for i, s in enumerate(up)}
    Lp = np.zeros((len(up), len(states)))
    for col, s in enumerate(states):
        for p in range(lz_max):
            c = np.sqrt((lz_max - p) * (p + 1.0))
            t, sg = jw_apply(s, [("c", p + 1), ("a", p)])
            if t is not None:
                Lp[idx_up[t], col] += sg * c
    lz = total_lz / 2.0
    return Lp.T @ Lp + (lz * lz + lz) * np.eye(len(states))


def dense_from_factors(n, total_lz, lz_max, f, off, shift):
    states = brute_sector(n, total_lz, lz_max)
    index = {s: i for i, s in enumerate(states)}
    H = shift * np.eye(len(states))
    for n1 in range(lz_max + 1):
        for n2 in range(n1):
            a = n1 * (n1 - 1) // 2 + n2
            S = n1 + n2
            lo = S // 2 + 1
            for m3 in range(lo, min(S, lz_max) + 1):
                F = f[off[a] + m3 - lo]
                if F == 0.0:
                    continue
                for col, s in enumerate(states):
                    t, sg = jw_apply(s, [("c", m3), ("c", S - m3), ("a", n1), ("a", n2)])
                    if t is not None:
                        H[index[t], col] += sg * F
    return H


@pytest.mark.parametrize("n,total_lz,lz_max", [(4, 0, 8), (4, 2, 8), (3, 1, 7), (5, 3, 9)])
def test_l2_factors_reproduce_ladder_operator(n, total_lz, lz_max):
    from diagham_amd import capi
    f, off, shift = capi.sphere_l2_factors(lz_max, n, total_lz, 1.0)
    H = dense_from_factors(n, total_lz, lz_max, f, off, shift)
    L2 = ladder_l2(n, total_lz, lz_max)
    assert np.max(np.abs(H - L2)) < 1e-12 * max(1.0, np.max(np.abs(L2)))


@pytest.mark.parametrize("n,total_lz,lz_max", [(4, 0, 8), (5, 1, 9)])
def test_l2_spectrum_is_l_times_l_plus_one(n, total_lz, lz_max):
    """Eigenvalues L(L+1) for L >= |Lz| (L integer or half-integer like Lz), each with
    multiplicity dim(Lz=L) - dim(Lz=L+1)."""
    from diagham_amd import capi
    f, off, shift = capi.sphere_l2_factors(lz_max, n, total_lz, 1.0)
    ev = np.sort(np.linalg.eigvalsh(dense_from_factors(n, total_lz, lz_max, f, off, shift)))
    top = n * (lz_max - n + 1)
    expect = []
    for twice_l in range(abs(total_lz), top + 1, 2):
        mult = capi.fermion_sphere_dimension(n, twice_l, lz_max) - (
            capi.fermion_sphere_dimension(n, twice_l + 2, lz_max) if twice_l + 2 <= top else 0)
        L = twice_l / 2.0
        expect += [L * (L + 1)] * mult
    assert len(expect) == ev.size
    assert np.max(np.abs(ev - np.sort(expect))) < 1e-10


def test_l2_factors_are_symmetric_and_scale():
    from diagham_amd import capi
    lz = 11
    f1, off, s1 = capi.sphere_l2_factors(lz, 6, 0, 1.0)
    f3, _, s3 = capi.sphere_l2_factors(lz, 6, 0, 3.0)
    assert np.allclose(f3, 3.0 * f1, rtol=0, atol=1e-13) and s3 == 3.0 * s1
    table = {}
    for n1 in range(lz + 1):
        for n2 in range(n1):
            S, lo = n1 + n2, (n1 + n2) // 2 + 1
            for m3 in range(lo, min(S, lz) + 1):
                table[(n1, n2, m3)] = f1[off[n1 * (n1 - 1) // 2 + n2] + m3 - lo]
    for (n1, n2, m3), v in table.items():
        m4 = n1 + n2 - m3
        assert abs(v - table[(m3, m4, n1)]) < 1e-13
