"""Three-body terms on the sphere (ProdA / ProdAd with explicit per-sum factor matrices): the
gfx950 gather kernel against the oracle's restatement of the reference's scatter loop."""
import itertools

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

N, TLZ, LZ = 6, 0, 13


@pytest.fixture(scope="module")
def tables():
    counts, tuples = O.nbody_tuples(LZ + 1, 3)
    F = O.random_nbody_factors(counts, 21)
    return counts, tuples, F


@pytest.fixture(scope="module")
def problem(arch, tables):
    from diagham_amd import capi
    counts, tuples, F = tables
    p = capi.Problem.sphere_nbody(arch, N, TLZ, LZ, 3, counts, tuples, F)
    yield p
    p.close()


@pytest.mark.parametrize("k,lpr", [(1, 1), (1, 4), (4, 1), (4, 8), (8, 2)])
def test_three_body_products_match_oracle(problem, tables, k, lpr):
    counts, tuples, F = tables
    s = O.FermionSphere(N, TLZ, LZ)
    assert problem.dim == s.dim
    assert np.array_equal(problem.states(), s.states())
    problem.set_launch(lanes_per_row=lpr, block_size=512)
    X = np.stack([O.splitmix(40 + v, s.dim) for v in range(k)], axis=1)
    Y = problem.hv(X.reshape(-1), k).reshape(s.dim, k)
    for v in range(k):
        y_ref = O.nbody_hv(s, 3, counts, tuples, F, X[:, v])
        assert np.linalg.norm(Y[:, v] - y_ref) <= 1e-12 * np.linalg.norm(y_ref), (k, lpr, v)


@pytest.mark.parametrize("lpr", [1, 4])
def test_three_body_products_larger_sector(arch, lpr):
    from diagham_amd import capi
    n, tlz, lz = 8, 2, 17
    counts, tuples = O.nbody_tuples(lz + 1, 3)
    F = O.random_nbody_factors(counts, 5)
    p = capi.Problem.sphere_nbody(arch, n, tlz, lz, 3, counts, tuples, F)
    p.set_launch(lanes_per_row=lpr, block_size=256)
    s = O.FermionSphere(n, tlz, lz)
    x = O.splitmix(9, s.dim)
    y = p.hv(x, 1)
    y_ref = O.nbody_hv(s, 3, counts, tuples, F, x)
    assert np.linalg.norm(y - y_ref) <= 1e-12 * np.linalg.norm(y_ref)
    p.close()


def test_three_body_element_count(problem, tables):
    counts, tuples, F = tables
    s = O.FermionSphere(N, TLZ, LZ)
    per_sum = {}
    t = tuples.reshape(-1, 3)
    pos = 0
    for S, c in enumerate(counts):
        per_sum[S] = [sum(1 << x for x in t[pos + l]) for l in range(c)]
        pos += c
    expect = 0
    for st in s.states():
        st = int(st)
        occ = [i for i in range(LZ + 1) if (st >> i) & 1]
        for tri in itertools.combinations(occ, 3):
            rest = st & ~sum(1 << x for x in tri)
            expect += sum(1 for m in per_sum[sum(tri)] if not (rest & m))
    assert problem.count_elements() == expect


@pytest.mark.parametrize("block", [1, 4])
def test_three_body_lanczos_matches_dense(problem, tables, block):
    counts, tuples, F = tables
    s = O.FermionSphere(N, TLZ, LZ)
    H = np.zeros((s.dim, s.dim))
    for c in range(s.dim):
        e = np.zeros(s.dim)
        e[c] = 1.0
        H[:, c] = O.nbody_hv(s, 3, counts, tuples, F, e)
    ref = np.linalg.eigvalsh(H)[:6]
    problem.set_launch(lanes_per_row=4, block_size=512)
    r = problem.lanczos(6, block_size=block)
    assert r["converged"]
    assert np.max(np.abs(r["eigenvalues"] - ref)) < 1e-10 * max(1.0, np.max(np.abs(ref)))


def test_asymmetric_three_body_factors_are_refused(arch, tables):
    from diagham_amd import capi
    counts, tuples, F = tables
    bad = F.copy()
    S = int(np.argmax(counts))  # a sum with several tuples
    off = int(np.sum(counts[:S] ** 2))
    bad[off + 1] += 0.5  # F_S[0][1] only
    with pytest.raises(capi.DHError, match="symmetric"):
        capi.Problem.sphere_nbody(arch, N, TLZ, LZ, 3, counts, tuples, bad)
