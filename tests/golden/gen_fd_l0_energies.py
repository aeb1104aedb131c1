"""Writes tests/golden/fd_n10_2s27_l0.json: the two lowest L = 0 energies of the differentiation
workload, and their central differences, for the finite-difference check of the adjoint kernel.

System: N = 10 fermions, 2S = 27, Lz = 0 sector (246,448 states), two-body interaction from
pseudopotentials V_m (the same draw as tests/golden/fd_n10_2s27.json: Coulomb(2S=27) rounded to 14
digits, the odd V_1..V_11 scaled by 1 + 0.01 g_m with g = numpy.random.default_rng(4).standard_normal(6)).
"Lz = 0 and L = 0 lowest states" (the states the reference selects with --l2-factor / --get-lvalue,
FQHESphereFermionsTwoBodyGeneric.cc:79-81, :170; ParticleOnSphereL2Hamiltonian.cc:240): the lowest
NBR_EIGEN eigenpairs of H in Lz = 0 are labelled by their total angular momentum, and the two lowest
with L = 0 are kept (H conserves L, so its eigenvectors in Lz = 0 have a definite L).

Here H x is the oracle's product, with every row's elements stored once (or_fs_csr_create,
or_csr_hv).  L is read from <L^2> = |L^+ psi|^2 (Lz = 0), with
L^+ = sum_p sqrt((2S - p)(p + 1)) a+_{p+1} a_p built independently from Jordan-Wigner ladder operators
(adjacent hops carry no sign) as a sparse map from the Lz = 0 to the Lz = 1 sector of the oracle's
bases.  The eigenpairs come from a full-reorthogonalisation Lanczos (three-term recurrence, then two
classical Gram-Schmidt passes over all previous vectors) with the strong stopping rule of
FullReorthogonalizedLanczosAlgorithm.cc:325 at 1e-14 on all NBR_EIGEN values.

Run from the repository root:  python tests/golden/gen_fd_l0_energies.py
"""
import json
import os
import sys
import time

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402

N, LZ_MAX = 10, 27
FREE_M = [1, 3, 5, 7, 9, 11]
H = 1e-5
SEED = 4
NBR_EIGEN = 14  # the second L = 0 state is the 14th state of the sector


def pseudopotentials():
    v = O.round14(O.coulomb_pseudopotentials(LZ_MAX))
    g = np.random.default_rng(SEED).standard_normal(len(FREE_M))
    for m, gm in zip(FREE_M, g):
        v[m] *= 1.0 + 0.01 * gm
    return v


def ladder_up(s0, s1):
    """L^+ from the Lz = 0 sector (s0) to the Lz = 1 sector (s1) as a sparse matrix"""
    st0 = s0.states().astype(np.int64)
    st1 = s1.states().astype(np.int64)
    order = np.argsort(st1)
    rows, cols, vals = [], [], []
    for p in range(LZ_MAX):
        m = (((st0 >> p) & 1) == 1) & (((st0 >> (p + 1)) & 1) == 0)
        src = np.nonzero(m)[0]
        t = st0[src] ^ (3 << p)
        idx = order[np.searchsorted(st1[order], t)]
        assert np.array_equal(st1[idx], t)
        rows.append(idx)
        cols.append(src)
        vals.append(np.full(src.size, np.sqrt((LZ_MAX - p) * (p + 1.0))))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(s1.dim, s0.dim))


def lowest(apply, dim, nev, precision=1e-14, max_iter=1500):
    """nev lowest eigenvalues and Ritz vectors of the symmetric operator `apply`"""
    V = np.zeros((max_iter + 1, dim))
    v = O.splitmix(0x5EED0000, dim)
    V[0] = v / np.linalg.norm(v)
    alpha, beta = [], []
    prev = None
    for j in range(max_iter):
        w = apply(V[j])
        a = V[j] @ w
        w -= a * V[j]
        if j > 0:
            w -= beta[-1] * V[j - 1]
        for _ in range(2):
            w -= V[:j + 1].T @ (V[:j + 1] @ w)
        b = np.linalg.norm(w)
        alpha.append(a)
        V[j + 1] = w / b
        if j + 1 >= nev + 1 and (j + 1) % 5 == 0:
            T = np.diag(alpha) + np.diag(beta, 1) + np.diag(beta, -1)
            ev = np.linalg.eigvalsh(T)[:nev]
            if prev is not None and np.all(np.abs(ev - prev) <= precision * np.abs(ev)):
                w_, U = np.linalg.eigh(T)
                return ev, (V[:j + 1].T @ U[:, :nev]).T, j + 1
            prev = ev
        beta.append(b)
    raise RuntimeError("Lanczos did not converge")


def energies(space, lp, v, threads):
    """the two lowest L = 0 eigenvalues of H(v), the L of the NBR_EIGEN lowest states, iterations"""
    rows = O.StoredRows(space, O.Factors(LZ_MAX, v), threads)
    ev, vecs, it = lowest(rows.hv, space.dim, NBR_EIGEN)
    labels = []
    for x in vecs:
        l2 = np.linalg.norm(lp @ x) ** 2 / (x @ x)
        L = (-1.0 + np.sqrt(1.0 + 4.0 * l2)) / 2.0
        assert abs(L - round(L)) < 1e-6, ("no definite L", L)
        labels.append(int(round(L)))
    zero = [i for i, L in enumerate(labels) if L == 0]
    assert len(zero) >= 2, ("fewer than two L = 0 states among the lowest", NBR_EIGEN, labels)
    return float(ev[zero[0]]), float(ev[zero[1]]), labels, it


def main():
    threads = int(os.environ.get("OMP_NUM_THREADS", "8"))
    s0 = O.FermionSphere(N, 0, LZ_MAX)
    s1 = O.FermionSphere(N, 2, LZ_MAX)
    lp = ladder_up(s0, s1)
    v0 = pseudopotentials()
    t = time.time()
    e0, e1, labels, it = energies(s0, lp, v0, threads)
    print("nominal", e0, e1, labels, it, "%.1fs" % (time.time() - t), flush=True)
    out = {"nbr_fermions": N, "lz_max": LZ_MAX, "total_lz": 0, "free_m": FREE_M, "h": H, "seed": SEED,
           "nbr_eigen": NBR_EIGEN, "total_l_of_lowest": labels, "pseudopotentials": v0.tolist(), "E0": e0,
           "E1": e1, "plus": {}, "minus": {}, "dE0_dV": [], "dGap_dV": []}
    for m in FREE_M:
        vp, vm = v0.copy(), v0.copy()
        vp[m] += H
        vm[m] -= H
        ep = energies(s0, lp, vp, threads)[:2]
        em = energies(s0, lp, vm, threads)[:2]
        out["plus"][str(m)] = ep
        out["minus"][str(m)] = em
        out["dE0_dV"].append((ep[0] - em[0]) / (2 * H))
        out["dGap_dV"].append(((ep[1] - ep[0]) - (em[1] - em[0])) / (2 * H))
        print("m", m, out["dE0_dV"][-1], out["dGap_dV"][-1], flush=True)
    with open(os.path.join(HERE, "fd_n10_2s27_l0.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
