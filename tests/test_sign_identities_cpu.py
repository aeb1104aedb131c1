"""Index and sign identities the device kernels rely on, checked exhaustively on the CPU: the row
ranges of the two-rows-per-lane dot products, and the fermionic-sign identities of the product
kernels against the reference's form of the sign (the parity of the occupied orbitals at and above each operator,
FermionOnSphere::AdAdAA, FQHE/src/HilbertSpace/FermionOnSphere.cc:468, and the AA / AdAd pair of
FermionOnSquareLatticeMomentumSpace).

Sphere (dh_sphere_row_free): creating m3 > m4 on the intermediate state t costs
popc(t >> m4) + popc(t >> m3); the kernel takes popc(t & ((1 << m3) - (1 << m4))), the occupied
orbitals in [m4, m3), which has the same parity.

Lattice (lattice_hv_kernel, free-pair rows): AdAd(m1, m2) with m1 < m2 creates m2 first, then m1:
popc(t >> m2) + popc((t | 1 << m2) >> m1); the kernel takes 1 + popc(t & ((1 << m2) - (1 << m1))).
"""
import itertools

import numpy as np


def popc(v):
    return bin(v).count("1")


def test_sphere_created_pair_sign():
    n = 12
    for t in range(1 << n):
        free = [m for m in range(n) if not (t >> m) & 1]
        for m4, m3 in itertools.combinations(free, 2):  # m4 < m3
            reference = popc(t >> m4) + popc(t >> m3)
            kernel = popc(t & ((1 << m3) - (1 << m4)))
            assert (reference - kernel) % 2 == 0, (t, m3, m4)


def test_lattice_created_pair_sign():
    n = 12
    for t in range(1 << n):
        free = [m for m in range(n) if not (t >> m) & 1]
        for m1, m2 in itertools.combinations(free, 2):  # m1 < m2
            reference = popc(t >> m2) + popc((t | (1 << m2)) >> m1)
            kernel = 1 + popc(t & ((1 << m2) - (1 << m1)))
            assert (reference - kernel) % 2 == 0, (t, m1, m2)


def test_negation_by_sign_bit_is_negation():
    """dh_negate_if flips the sign bit of the factor: the same double as -c, zeros included."""
    rng = np.random.default_rng(3)
    c = np.concatenate([rng.standard_normal(1000), [0.0, -0.0, 1e-310, np.inf]])
    flipped = (c.view(np.uint64) ^ np.uint64(1 << 63)).view(np.float64)
    assert np.array_equal(flipped.view(np.uint64), (-c).view(np.uint64))


def test_row_pair_ranges_cover_every_row_once():
    """The k = 1 dot products with two rows per lane (dh_dots_rows_pairs): the aligned pairs
    (2q, 2q + 1) for q in [(b + 1) / 2, e / 2), an odd first row b on lane 0 and an odd last row e - 1
    on lane 1, cover every row of [b, e) exactly once, for every range of up to 40 rows."""
    for b in range(0, 40):
        for e in range(b, 41):
            rows = []
            pb, pe = (b + 1) >> 1, e >> 1
            for q in range(pb, pe):
                rows += [2 * q, 2 * q + 1]
            if (b & 1) and b < e:
                rows.append(b)
            if (e & 1) and e - 1 >= b:
                rows.append(e - 1)
            assert sorted(rows) == list(range(b, e)), (b, e, rows)
