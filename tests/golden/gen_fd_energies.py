"""Writes tests/golden/fd_n10_2s27.json: oracle Lanczos energies for the finite-difference
check of the adjoint kernel.

System: N = 10 fermions, 2S = 27, Lz = 0 sector (246,448 states), two-body interaction
from pseudopotentials.  V_m = Coulomb(2S=27) rounded to 14 digits for every odd m, with the
free coefficients V_1, V_3, ..., V_11 drawn around their Coulomb values:
V_m *= 1 + 0.01 * g_m,  g = numpy.random.default_rng(4).standard_normal(6).
For every free m the two lowest eigenvalues E0, E1 are computed at V +- h e_m (h = 1e-5)
with the oracle's full-reorthogonalisation Lanczos (strong convergence at 1e-14), and the
central differences dE0/dV_m, d(E1-E0)/dV_m are stored.

Run from the repository root:  python tests/golden/gen_fd_energies.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402

N, LZ_MAX, TOTAL_LZ = 10, 27, 0
FREE_M = [1, 3, 5, 7, 9, 11]
H = 1e-5
SEED = 4


def pseudopotentials():
    v = O.round14(O.coulomb_pseudopotentials(LZ_MAX))
    g = np.random.default_rng(SEED).standard_normal(len(FREE_M))
    for m, gm in zip(FREE_M, g):
        v[m] *= 1.0 + 0.01 * gm
    return v


def energies(space, v, threads):
    f = O.Factors(LZ_MAX, v)
    ev, it = O.fs_lanczos(space, f, 2, threads=threads, seed=0x5EED0000)
    assert it > 0, "oracle Lanczos did not converge"
    return float(ev[0]), float(ev[1])


def main():
    threads = int(os.environ.get("OMP_NUM_THREADS", "8"))
    space = O.FermionSphere(N, TOTAL_LZ, LZ_MAX)
    v0 = pseudopotentials()
    t = time.time()
    e0, e1 = energies(space, v0, threads)
    print("nominal", e0, e1, "%.1fs" % (time.time() - t), flush=True)
    out = {"nbr_fermions": N, "lz_max": LZ_MAX, "total_lz": TOTAL_LZ, "free_m": FREE_M, "h": H, "seed": SEED,
           "pseudopotentials": v0.tolist(), "E0": e0, "E1": e1, "plus": {}, "minus": {},
           "dE0_dV": [], "dGap_dV": []}
    for m in FREE_M:
        vp, vm = v0.copy(), v0.copy()
        vp[m] += H
        vm[m] -= H
        ep = energies(space, vp, threads)
        em = energies(space, vm, threads)
        out["plus"][str(m)] = ep
        out["minus"][str(m)] = em
        out["dE0_dV"].append((ep[0] - em[0]) / (2 * H))
        out["dGap_dV"].append(((ep[1] - ep[0]) - (em[1] - em[0])) / (2 * H))
        print("m", m, out["dE0_dV"][-1], out["dGap_dV"][-1], flush=True)
    with open(os.path.join(HERE, "fd_n10_2s27.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
