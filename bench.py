"""Benchmark of the on-the-fly H.v + Lanczos path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Headline workload: spin-1/2 Heisenberg chain, L = 28 sites, periodic, J = 1, Sz = 0
(40,116,600 states).  One step = one H.v over the whole basis (the one-vector product).
The JSON line carries matrix elements / s, the roofline of that kernel (algorithmic HBM
bytes per launch from HIP events on the kernel's stream), the measured HBM traffic of that kernel
(rocprofv3 counters kept under profiles/) as a rate, and beside it:
  lanczos           time to the 4 lowest eigenvalues of that chain; block vs plain to the 10
                    lowest of N = 8 fermions at 2S = 21; and "all_sectors": time to the 10 lowest
                    eigenvalues of every Lz >= 0 sector of N = 10 fermions at 2S = 27, the plain
                    one-vector Lanczos sector after sector vs the batched block Lanczos
                    (plain, batched k = 1, batched block k = 8), the same for the 133 sectors of
                    N = 12 at 2S = 33 (plain vs batched block k = 8) and for the 30 (kx, ky) sectors
                    of N = 10 on 6 x 5 checkerboard sites (plain vs batched k = 1 and k = 4), and the
                    sector autotuner's choice for N = 10 at 2S = 27
  index_lookup      the headline product with the LUT-seeded binary search instead of the rank
  adjoint           dE0/dV_m of N = 10 at 2S = 27, Lz = 0 in one pass, beside one product
  batched_roofline  one k = 8 product over every Lz >= 0 sector of N = 12 fermions at 2S = 33
                    in one persistent launch (the batching workload's kernel)
  cpu_baseline      the oracle's CPU rate on a bounded sample of the headline product, and its
                    sphere and lattice products and a short Lanczos solve
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "H·v matrix elements/s + Lanczos time-to-10-eigvals, block vs plain, %HBM peak"
# the all-sector legs of N=12 2S=33 (about 15 minutes), recorded by `bench.py --n12 ...` runs
N12_RECORD = "sectors_n12_2s33_time_to_eigenvalues.json"
HBM_PEAK = 8.0e12


class HipTimer:
    """hipEvent pair on a given stream (the stream the kernels are launched on)."""

    def __init__(self, stream):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.stream = stream
        self.a, self.b = ctypes.c_void_p(), ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(self.a)) == 0
        assert self.hip.hipEventCreate(ctypes.byref(self.b)) == 0

    def start(self):
        assert self.hip.hipEventRecord(self.a, self.stream) == 0

    def stop_ms(self):
        assert self.hip.hipEventRecord(self.b, self.stream) == 0
        assert self.hip.hipEventSynchronize(self.b) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), self.a, self.b) == 0
        return float(ms.value)


def make_stream():
    hip = ctypes.CDLL("libamdhip64.so")
    s = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(s)) == 0  # default priority
    return s


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def spin_hv_bench(capi, arch, stream, L, steps, warmup, k=1, index_mode=None):
    p = capi.Problem.spin_chain(arch, L, 0, 1.0, 1.0, True)
    # 1024-thread workgroups: the fastest measured shape
    p.set_launch(1, 1024, capi.INDEX_RANK if index_mode is None else index_mode, 1)
    nnz = p.count_elements()
    x = capi.DeviceBuffer(arch, p.dim * k)
    y = capi.DeviceBuffer(arch, p.dim * k)
    capi.gpu().dh_fill_random(arch.ctx, x.ptr, p.dim * k, 0x5EED0001)
    for _ in range(warmup):
        p.hv_device(x.ptr, y.ptr, k)
    arch.sync()
    timer = HipTimer(stream)
    t0 = time.perf_counter()
    timer.start()
    for _ in range(steps):
        p.hv_device(x.ptr, y.ptr, k)
    ms = timer.stop_ms()
    arch.sync()
    wall = time.perf_counter() - t0
    x.free()
    y.free()
    out = {"dim": p.dim, "nnz": nnz, "ms_per_step": ms / steps, "wall_s": wall}
    p.close()
    return out


def copy_calibration(capi, arch, stream, n=1 << 27, reps=10):
    """Streaming copy of n doubles (the HBM-traffic calibration kernel): GB/s."""
    x = capi.DeviceBuffer(arch, n)
    y = capi.DeviceBuffer(arch, n)
    capi.gpu().dh_fill_random(arch.ctx, x.ptr, n, 3)
    for _ in range(2):
        capi.gpu().dh_copy(arch.ctx, y.ptr, x.ptr, n)
    timer = HipTimer(stream)
    timer.start()
    for _ in range(reps):
        capi.gpu().dh_copy(arch.ctx, y.ptr, x.ptr, n)
    ms = timer.stop_ms() / reps
    x.free()
    y.free()
    return {"doubles": n, "ms": ms, "GBps": 16.0 * n / (ms * 1e-3) / 1e9}


def progress(msg):
    """One progress line per finished leg, on stderr (stdout carries only the JSON line)."""
    print("[bench %.1f s] %s" % (time.perf_counter() - _T0, msg), file=sys.stderr, flush=True)


_T0 = time.perf_counter()


def lanczos_runs(capi, arch, report=None):
    res = {}
    # time to the 4 lowest eigenvalues of the L = 28 chain (plain, full reorthogonalisation)
    p = capi.Problem.spin_chain(arch, 28, 0)
    r = p.lanczos(4, block_size=1, max_iter=400)
    res["heisenberg_L28_4_eigenvalues"] = {"seconds": r["seconds"], "products": r["products"],
                                           "eigenvalues": r["eigenvalues"].tolist(), "converged": r["converged"]}
    p.close()
    if report:
        report("L=28 Lanczos: %.2f s" % r["seconds"])
    # block vs plain to the 10 lowest eigenvalues of N = 8 fermions at 2S = 21, Lz = 0
    pp = capi.coulomb_pseudopotentials(21, True)
    p = capi.Problem.sphere(arch, 8, 0, 21, pp)
    out = {}
    for block in (1, 8):
        r = p.lanczos(10, block_size=block)
        out["k%d" % block] = {"seconds": r["seconds"], "products": r["products"], "blocks": r["blocks"],
                              "e0": float(r["eigenvalues"][0]), "e9": float(r["eigenvalues"][9])}
    out["speedup_block_over_plain"] = out["k1"]["seconds"] / out["k8"]["seconds"]
    res["fermions_n8_2s21_10_eigenvalues"] = out
    p.close()
    return res


def sector_runs(capi, arch, report=None, n=10, lz=27, nev=10, legs=(("plain_sequential_k1", 1, False),
                                                                     ("batched_k1", 1, True),
                                                                     ("batched_block_k8", 8, True)), dump=None):
    """Time to the nev lowest eigenvalues of every Lz >= 0 sector of n Coulomb fermions at flux lz:
    the plain one-vector Lanczos run sector after sector (the reference's loop,
    FQHESphereFermionsTwoBodyGeneric.cc:270) against the batched Lanczos that advances all sectors
    together, with the waves the device memory needed, the device memory peak and the largest
    eigenvalue difference from the plain run."""
    pp = capi.coulomb_pseudopotentials(lz, True)
    lzs = capi.sphere_sectors(n, lz)
    dims = [capi.fermion_sphere_dimension(n, s, lz) for s in lzs]
    runs = {}
    for name, k, batched in legs:
        arch.memory_peak(reset=True)
        r = capi.sectors_solve(arch, n, lz, lzs, pp, nev, k, batched=batched, allow_unconverged=True)
        r["memory_peak"] = arch.memory_peak()
        runs[name] = r
        if report:
            report("sectors N=%d 2S=%d, %s: %.2f s, %d waves" % (n, lz, name, r["seconds"], r["waves"]))
    if dump:  # every leg's eigenvalues, to compare legs run in separate processes
        with open(dump, "w") as f:
            json.dump({name: r["eigenvalues"].tolist() for name, r in runs.items()}, f)
    return sector_summary("fermions_sphere_n%d_2s%d_coulomb_all_Lz_ge_0_sectors_%d_eigenvalues" % (n, lz, nev),
                          runs, legs[0][0], len(lzs), dims)


def sector_summary(workload, runs, plain, nbr_sectors, dims):
    """Per leg: time, products, converged sectors, waves, device memory peak, and the largest
    eigenvalue difference from the first leg (the plain run when it is among the legs)."""
    ref = runs[plain]["eigenvalues"]
    out = {"workload": workload, "sectors": nbr_sectors, "states": int(sum(dims)), "largest_sector": int(max(dims))}
    for name, r in runs.items():
        out[name] = {"seconds": r["seconds"], "products": int(np.sum(r["products"])),
                     "converged": int(np.sum(r["converged"])), "waves": r.get("waves", 0),
                     "device_memory_peak_bytes": r.get("memory_peak"),
                     "max_abs_eigenvalue_difference_vs_plain": float(np.max(np.abs(r["eigenvalues"] - ref)))}
    t0 = runs[plain]["seconds"]
    if plain.startswith("plain"):
        for name in runs:
            if name != plain:
                out[name]["speedup_vs_plain"] = t0 / runs[name]["seconds"]
    return out


def lattice_runs(capi, arch, report=None, n=10, nx=6, ny=5, nev=5, block=4):
    """Time to the nev lowest eigenvalues of all nx * ny (kx, ky) sectors of n fermions in the
    projected checkerboard band: the plain one-vector complex Lanczos sector after sector (the loop
    of FCICheckerboardLatticeModel.cc:412) against the batched complex block Lanczos over all
    sectors."""
    kx = [x for x in range(nx) for y in range(ny)]
    ky = [y for x in range(nx) for y in range(ny)]
    runs = {}
    arch.memory_peak(reset=True)
    t0 = time.perf_counter()
    ev, prods, conv = [], [], []
    for x, y in zip(kx, ky):
        r = capi.lattice_sectors_solve(arch, n, nx, ny, [x], [y], nev, 1, allow_unconverged=True)
        ev.append(r["eigenvalues"][0])
        prods.append(int(r["products"][0]))
        conv.append(bool(r["converged"][0]))
    runs["plain_sequential_k1"] = {"eigenvalues": np.array(ev), "products": np.array(prods),
                                   "converged": np.array(conv), "seconds": time.perf_counter() - t0, "waves": 0,
                                   "memory_peak": arch.memory_peak()}
    if report:
        report("lattice %dx%d N=%d, plain_sequential_k1: %.2f s" % (nx, ny, n, runs["plain_sequential_k1"]["seconds"]))
    for name, k in (("batched_k1", 1), ("batched_block_k%d" % block, block)):
        arch.memory_peak(reset=True)
        r = capi.lattice_sectors_solve(arch, n, nx, ny, kx, ky, nev, k, allow_unconverged=True)
        r["memory_peak"] = arch.memory_peak()
        runs[name] = r
        if report:
            report("lattice %dx%d N=%d, %s: %.2f s, %d waves" % (nx, ny, n, name, r["seconds"], r["waves"]))
    dims = capi.momentum_sector_dims(arch, n, nx, ny, kx, ky)
    return sector_summary("checkerboard_%dx%d_n%d_t1_1_t2_0.5_tpp_-0.5_V1_flat_all_%d_sectors_%d_eigenvalues" %
                          (nx, ny, n, nx * ny, nev), runs, "plain_sequential_k1", nx * ny, dims)


def batched_hv_roofline(capi, arch, stream, n=12, lz=33, k=8, reps=3):
    """One persistent product of k vectors over every Lz >= 0 sector of n fermions at flux lz
    (the batching workload's kernel), timed with HIP events on the bench stream."""
    pp = capi.coulomb_pseudopotentials(lz, True)
    lzs = capi.sphere_sectors(n, lz)
    S = capi.SphereSectorSet(arch, n, lz, lzs, pp)
    # two lanes per row, 1024-thread workgroups: the sector solver's own shape for k >= 4, the fastest
    # of the committed sweep (profiles/launch_shape_sweeps.md)
    S.set_launch(2, 1024)
    nnz = int(S.count_elements().sum())
    x = capi.DeviceBuffer(arch, S.total * k)
    y = capi.DeviceBuffer(arch, S.total * k)
    capi.gpu().dh_fill_random(arch.ctx, x.ptr, S.total * k, 0x5EED0003)
    S.hv_device(x.ptr, y.ptr, k)
    arch.sync()
    timer = HipTimer(stream)
    timer.start()
    for _ in range(reps):
        S.hv_device(x.ptr, y.ptr, k)
    ms = timer.stop_ms() / reps
    x.free()
    y.free()
    rows = S.total
    S.close()
    alg = rows * (8 + 8 * k) + nnz * 8 * k
    achieved = alg / (ms * 1e-3) / 1e9
    traffic = load_traffic("traffic_sectors_hv_n%d_2s%d_k%d.json" % (n, lz, k))
    return {"workload": "fermions_sphere_n%d_2s%d_all_Lz_ge_0_sectors_one_product_k%d" % (n, lz, k),
            "kernel": "sectors_hv_kernel<%d, 2, true, false, false>" % k, "sectors": len(lzs), "rows": rows,
            "launch": {"lanes_per_row": 2, "block_size": 1024, "row_routine": "range", "low_rank_table": "LDS, 16-bit",
                       "chosen_by": "profiles/launch_shape_sweeps.md"},
            "matrix_elements": nnz, "ms": ms, "element_vector_products_per_s": nnz * k / (ms * 1e-3),
            "algorithmic_bytes": alg, "achieved_GBps": achieved, "frac_of_peak": achieved / (HBM_PEAK / 1e9),
            "traffic": traffic and traffic["bytes_per_launch"],
            "measured_traffic_GBps": traffic and traffic["bytes_per_launch"] / (ms * 1e-3) / 1e9}


def adjoint_leg(capi, arch, stream, n=10, lz=27, reps=10):
    """Gradient of the ground-state energy of N = 10 fermions at 2S = 27, Lz = 0 with respect to every
    pseudopotential: dh_sphere_adjoint forms all channels' bilinears <psi|dH/dV_m|psi> in one pass
    over the (N-2)-particle states (the channel interaction is of rank one in the pairs, so the
    bilinear factorises: one gather per row and annihilated pair), timed beside one product of the
    same Hamiltonian (both on the bench stream)."""
    pp = capi.coulomb_pseudopotentials(lz, True)
    p = capi.Problem.sphere(arch, n, 0, lz, pp)
    psi = capi.DeviceBuffer(arch, p.dim)
    y = capi.DeviceBuffer(arch, p.dim)
    r = p.lanczos(1, d_vectors=psi.ptr, nbr_vectors=1)
    nnz = p.count_elements()
    timer = HipTimer(stream)
    g = p.adjoint_device(psi.ptr, psi.ptr)
    timer.start()
    for _ in range(reps):
        g = p.adjoint_device(psi.ptr, psi.ptr)
    ms_adj = timer.stop_ms() / reps
    p.hv_device(psi.ptr, y.ptr)
    arch.sync()
    timer.start()
    for _ in range(reps):
        p.hv_device(psi.ptr, y.ptr)
    ms_hv = timer.stop_ms() / reps
    psi.free()
    y.free()
    p.close()
    # one gather of psi per (row, annihilated pair): C(N, 2) per row (psi and phi are the same vector)
    alg = p.dim * (n * (n - 1) // 2) * 8
    return {"workload": "fermions_sphere_n%d_2s%d_lz0_ground_state_dE_dVm" % (n, lz), "dimension": p.dim,
            "channels": len(g), "matrix_elements": nnz, "e0": float(r["eigenvalues"][0]), "dE0_dV1": float(g[0]),
            "adjoint_ms": ms_adj,
            "product_ms": ms_hv, "adjoint_over_product": ms_adj / ms_hv,
            "algorithmic_bytes": alg, "achieved_GBps": alg / (ms_adj * 1e-3) / 1e9,
            "frac_of_peak": alg / (ms_adj * 1e-3) / HBM_PEAK}


def reorth_stream_leg(capi, arch, stream, L=28, nb=100, reps=3):
    """The two passes of a full reorthogonalisation step at the Krylov depth of the L = 28 Lanczos
    (100 basis vectors of the chain's 40,116,600 rows, 32 GB): dh_block_dots (C = V^T w, with its
    chunk reduction and the copy of C to the host) and dh_block_axpy (w -= V C), each timed with HIP
    events on the bench stream.  Bytes: the basis once plus w once (dots), plus w written (axpy)."""
    dim = math.comb(L, L // 2)
    V = capi.DeviceBuffer(arch, nb * dim)
    w = capi.DeviceBuffer(arch, dim)
    capi.gpu().dh_fill_random(arch.ctx, V.ptr, nb * dim, 0x5EED0009)
    capi.gpu().dh_fill_random(arch.ctx, w.ptr, dim, 0x5EED000A)
    out = np.zeros(nb)
    coef = np.full(nb, -1e-3)
    timer = HipTimer(stream)
    res = {}
    for name, call, extra in (("dots", lambda: capi.gpu().dh_block_dots(arch.ctx, V.ptr, nb, w.ptr, dim, 1,
                                                                         capi.dptr(out)), 1),
                              ("axpy", lambda: capi.gpu().dh_block_axpy(arch.ctx, V.ptr, nb, w.ptr, dim, 1,
                                                                         capi.dptr(coef)), 2)):
        assert call() == 0
        arch.sync()
        timer.start()
        for _ in range(reps):
            assert call() == 0
        ms = timer.stop_ms() / reps
        nbytes = 8.0 * dim * (nb + extra)
        res[name] = {"ms": ms, "bytes": nbytes, "GBps": nbytes / (ms * 1e-3) / 1e9,
                     "frac_of_peak": nbytes / (ms * 1e-3) / HBM_PEAK}
    V.free()
    w.free()
    return dict({"workload": "reorthogonalisation_L%d_%d_basis_vectors" % (L, nb), "rows": dim,
                 "basis_vectors": nb}, **res)


def sectors_autotune_leg(capi, arch, n=10, lz=27, nev=10, ks=(1, 2, 4, 8)):
    """dhh_sectors_autotune over every Lz >= 0 sector of N = 10 at 2S = 27: per k the fastest product
    launch on calibration sectors, scored by the time of a calibration solve with it."""
    pp = capi.coulomb_pseudopotentials(lz, True)
    t0 = time.perf_counter()
    rows, best = capi.sectors_autotune(arch, n, lz, capi.sphere_sectors(n, lz), pp, nev, ks=ks)
    return {"workload": "fermions_sphere_n%d_2s%d_all_Lz_ge_0_sectors_%d_eigenvalues" % (n, lz, nev),
            "seconds": time.perf_counter() - t0, "table": rows, "best": best}


def kernel_only(capi, arch, stream, name, steps, warmup):
    """One named product repeated `steps` times after `warmup` (for rocprofv3 runs): prints a JSON
    line with its mean time and rates."""
    k = int(name.split("_k")[1].split("_")[0]) if "_k" in name else 1
    timer = HipTimer(stream)
    cleanup = []
    free = "_free" in name  # the free-pair row routine for every k (DH_INDEX_RANK_FREE)
    rng = "_range" in name  # the range row routine for every k (DH_INDEX_RANK_RANGE)
    lpr = int(name.split("_lpr")[1].split("_")[0]) if "_lpr" in name else None
    bs = int(name.split("_bs")[1].split("_")[0]) if "_bs" in name and "_bsearch" not in name else None
    if name.startswith("sphere_n12_2s33_lz0"):
        p = capi.Problem.sphere(arch, 12, 0, 33, capi.coulomb_pseudopotentials(33, True))
        cleanup.append(p.close)
        bsearch = "_bsearch" in name
        mode = (capi.INDEX_RANK_FREE if free else capi.INDEX_RANK_RANGE if rng else
                capi.INDEX_BSEARCH if bsearch else capi.INDEX_RANK)
        if lpr or bs or free or rng or bsearch:  # otherwise the Hamiltonian's own choice per k
            p.set_launch(lpr or 2, bs or (256 if k <= 2 else 1024), mode, 1)
        rows, nnz = p.dim, p.count_elements()
        run = lambda dx, dy: p.hv_device(dx, dy, k)  # noqa: E731
        width = k
    elif name.startswith("sectors_n12_2s33"):
        S = capi.SphereSectorSet(arch, 12, 33, capi.sphere_sectors(12, 33), capi.coulomb_pseudopotentials(33, True))
        cleanup.append(S.close)
        S.set_launch(lpr or 2, bs or 1024) if k >= 4 else S.set_launch(lpr or 1, bs or 1024)
        if free or rng:
            S.set_row_mode(capi.INDEX_RANK_FREE if free else capi.INDEX_RANK_RANGE)
        rows, nnz = S.total, int(S.count_elements().sum())
        run = lambda dx, dy: S.hv_device(dx, dy, k)  # noqa: E731
        width = k
    elif name.startswith("lattice_6x5_n10"):
        from_lattice = capi.lattice_set(arch, 10, 6, 5, [x for x in range(6) for y in range(5)],
                                        [y for x in range(6) for y in range(5)])
        cleanup.append(from_lattice.close)
        from_lattice.set_launch(lpr or (2 if k <= 2 else 1), bs or (512 if k <= 2 else 256))
        if free or rng:
            arch.check(capi.gpu().dh_sector_set_row_mode(from_lattice.h, capi.INDEX_RANK_FREE if free
                                                         else capi.INDEX_RANK_RANGE), "row mode")
        rows, nnz = from_lattice.total, int(from_lattice.count_elements().sum())
        run = lambda dx, dy: from_lattice.hv_device(dx, dy, k)  # noqa: E731
        width = 2 * k
    else:
        raise SystemExit("unknown kernel workload " + name)
    x = capi.DeviceBuffer(arch, rows * width)
    y = capi.DeviceBuffer(arch, rows * width)
    capi.gpu().dh_fill_random(arch.ctx, x.ptr, rows * width, 0x5EED0007)
    for _ in range(warmup):
        run(x.ptr, y.ptr)
    arch.sync()
    timer.start()
    for _ in range(steps):
        run(x.ptr, y.ptr)
    ms = timer.stop_ms() / steps
    x.free()
    y.free()
    for f in cleanup:
        f()
    out = {"kernel_workload": name, "rows": rows, "k": k, "ms": ms,
           # the HBM-traffic calibration kernel in the same process, so that a counter run over this
           # command sees both (dh_copy_kernel: 2^27 doubles read and written per launch)
           "copy_calibration": copy_calibration(capi, arch, stream)}
    if nnz:
        alg = rows * (8 + 8 * width) + nnz * 8 * width
        out.update({"matrix_elements": nnz, "element_vector_products_per_s": nnz * k / (ms * 1e-3),
                    "algorithmic_GBps": alg / (ms * 1e-3) / 1e9})
    print(json.dumps(out))


def launch_sweep(capi, arch, stream, base, ks, reps=3):
    """Every launch shape (lanes per row 1, 2, 4 x workgroups of 256, 512, 1024 threads) of one
    product workload (sphere_n12_2s33_lz0, sectors_n12_2s33 or lattice_6x5_n10) for each k in ks:
    one JSON line per shape with its mean time over reps products (HIP events on the bench stream)."""
    timer = HipTimer(stream)
    if base == "sphere_n12_2s33_lz0":
        obj = capi.Problem.sphere(arch, 12, 0, 33, capi.coulomb_pseudopotentials(33, True))
        rows, width = obj.dim, 1
        shape = lambda lpr, bs: obj.set_launch(lpr, bs, capi.INDEX_RANK, 1)  # noqa: E731
    elif base == "sectors_n12_2s33":
        obj = capi.SphereSectorSet(arch, 12, 33, capi.sphere_sectors(12, 33), capi.coulomb_pseudopotentials(33, True))
        rows, width = obj.total, 1
        shape = obj.set_launch
    elif base == "lattice_6x5_n10":
        obj = capi.lattice_set(arch, 10, 6, 5, [x for x in range(6) for y in range(5)],
                               [y for x in range(6) for y in range(5)])
        rows, width = obj.total, 2
        shape = obj.set_launch
    else:
        raise SystemExit("unknown sweep workload " + base)
    for k in ks:
        x = capi.DeviceBuffer(arch, rows * width * k)
        y = capi.DeviceBuffer(arch, rows * width * k)
        capi.gpu().dh_fill_random(arch.ctx, x.ptr, rows * width * k, 0x5EED0007)
        for lpr in (1, 2, 4):
            for bs in (256, 512, 1024):
                shape(lpr, bs)
                obj.hv_device(x.ptr, y.ptr, k)
                arch.sync()
                timer.start()
                for _ in range(reps):
                    obj.hv_device(x.ptr, y.ptr, k)
                ms = timer.stop_ms() / reps
                print(json.dumps({"sweep": base, "k": k, "lanes_per_row": lpr, "block_size": bs, "ms": ms}))
                sys.stdout.flush()
        x.free()
        y.free()
    obj.close()


def cpu_threads():
    """Threads for the CPU legs: every core this process may run on, unless OMP_NUM_THREADS sets the
    host share (on a shared GPU host the share, not the machine, is what one GPU's job has)."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS")
    threads = min(affinity, int(share)) if share and share.isdigit() and int(share) > 0 else affinity
    return max(1, threads), {"nproc": os.cpu_count(), "affinity": affinity, "omp_num_threads": share}


def cpu_baseline(L=28, seconds_target=15.0):
    """Oracle (CPU restatement of SpinChainHamiltonian::LowLevelAddMultiply, reference thread
    split) on the host cores, over a bounded row sample of the same L = 28 product; beside it the
    oracle's sphere and lattice products and a short Lanczos solve."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    cores, host = cpu_threads()
    s = O.SpinChain(L, 0)
    J = np.ones(L)
    x = O.splitmix(1, s.dim)
    y = np.zeros(s.dim)
    rows = min(s.dim, 1 << 18)
    t0 = time.perf_counter()
    O.lib().or_spin_hv_rows(s.h, O.dptr(J), O.dptr(J), 1, O.dptr(x), O.dptr(y), 0, rows, cores)
    dt = time.perf_counter() - t0
    # grow the sample to about seconds_target of CPU work: more rows, then whole products repeated
    passes = 1
    while dt < seconds_target / 2 and (rows < s.dim or passes == 1):
        scale = seconds_target / max(dt, 1e-3)
        if rows < s.dim:
            rows = int(min(s.dim, rows * scale))
        else:
            passes = max(2, int(round(scale)))
        t0 = time.perf_counter()
        for _ in range(passes):
            y[:] = 0.0
            O.lib().or_spin_hv_rows(s.h, O.dptr(J), O.dptr(J), 1, O.dptr(x), O.dptr(y), 0, rows, cores)
        dt = time.perf_counter() - t0
    states = s.states()[:rows] & ((1 << L) - 1)
    rot = ((states >> 1) | ((states & 1) << (L - 1)))
    anti = np.unpackbits((states ^ rot).astype(">u4").view(np.uint8)).sum()
    elements = (int(anti) + rows) * passes  # off-diagonal flips + diagonal, per pass
    del s, x, y
    out = {"value": elements / dt, "unit": "matrix_elements/s", "cores": cores, "kind": "port",
           "sample": "oracle H.v over rows [0, %d) of the L=28 basis, %d pass(es) (%d matrix elements, %.1f s)" %
                     (rows, passes, elements, dt), "host": host}
    # sphere: one whole product of N = 10 fermions at 2S = 27, Lz = 0 (reference SMP split)
    sp = O.FermionSphere(10, 0, 27)
    fac = O.Factors(27, O.round14(O.coulomb_pseudopotentials(27)))
    xs = O.splitmix(2, sp.dim)
    t0 = time.perf_counter()
    O.fs_hv(sp, fac, xs, threads=cores)
    dt = time.perf_counter() - t0
    nnz = 57096228  # matrix elements of that product (dh_hamiltonian_count_elements)
    out["sphere_n10_2s27_hv"] = {"value": nnz / dt, "unit": "matrix_elements/s", "seconds": dt,
                                 "sample": "one oracle product over all 246,448 rows"}
    # lattice: rows [0, 100000) of sector (0, 0) of N = 10 on 6 x 5 sites, one complex vector
    u, _ = O.checkerboard_lower_band(6, 5, 1.0, 0.5, -0.5)
    ms = O.MomentumFermions(10, 6, 5, 0, 0)
    lf = O.LatticeFactors(6, 5, u)
    xl = O.splitmix(3, ms.dim) + 1j * O.splitmix(4, ms.dim)
    t0 = time.perf_counter()
    O.fm_hv_rows(ms, lf, xl, 0, 100000, threads=cores)
    dt = time.perf_counter() - t0
    out["lattice_6x5_n10_rows"] = {"value": 100000 / dt, "unit": "rows/s", "seconds": dt,
                                   "sample": "oracle gather rows [0, 100000) of sector (0, 0), one complex vector"}
    # time to the 10 lowest eigenvalues of N = 8 at 2S = 21 (oracle full-reorthogonalisation Lanczos)
    s8 = O.FermionSphere(8, 0, 21)
    f8 = O.Factors(21, O.round14(O.coulomb_pseudopotentials(21)))
    t0 = time.perf_counter()
    ev, it = O.fs_lanczos(s8, f8, 10, threads=cores)
    out["lanczos_n8_2s21_10_eigenvalues"] = {"seconds": time.perf_counter() - t0, "iterations": int(it)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--skip-lanczos", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-batched", action="store_true")
    ap.add_argument("--skip-n10", action="store_true", help="skip the all-sector legs of N=10 at 2S=27")
    ap.add_argument("--n12", default="",
                    help="run the all-sector legs of N=12 at 2S=33 (comma-separated: plain_sequential_k1, "
                         "batched_k1, batched_block_k8, or 'all'); they take about 15 minutes together, so by "
                         "default the line carries their recorded run from profiles/")
    ap.add_argument("--skip-lattice", action="store_true", help="skip the all-sector legs of N=10 on 6x5 sites")
    ap.add_argument("--n12-dump", default=None, help="file for every N=12 leg's eigenvalues (JSON)")
    ap.add_argument("--length", type=int, default=28)
    ap.add_argument("--kernel", default=None,
                    help="only repeat one product (sphere_n12_2s33_lz0_k1, sectors_n12_2s33_k8, "
                         "lattice_6x5_n10_k4, ...) for profiling")
    ap.add_argument("--sweep", default=None,
                    help="time every launch shape of one product workload (sphere_n12_2s33_lz0, "
                         "sectors_n12_2s33, lattice_6x5_n10) at the k of --sweep-k")
    ap.add_argument("--sweep-k", default="1", help="comma-separated block sizes k for --sweep")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one process, one MI355X); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from diagham_amd import capi
    arch = capi.Arch(0)
    stream = make_stream()
    arch.check(capi.gpu().dh_set_stream(arch.ctx, stream), "dh_set_stream")
    ncu, name = arch.device_info()

    if args.kernel:
        kernel_only(capi, arch, stream, args.kernel, args.steps, args.warmup)
        arch.close()
        return
    if args.sweep:
        launch_sweep(capi, arch, stream, args.sweep, [int(v) for v in args.sweep_k.split(",")])
        arch.close()
        return
    hv = spin_hv_bench(capi, arch, stream, args.length, args.steps, args.warmup, k=1)
    dim, nnz, ms = hv["dim"], hv["nnz"], hv["ms_per_step"]
    alg_bytes = dim * (8 + 8) + nnz * 8  # state word + y write per row, one x read per element
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    traffic = load_traffic("traffic_spin_hv_L%d.json" % args.length)
    calib = copy_calibration(capi, arch, stream)
    line = {
        "metric": METRIC,
        "value": nnz / (ms * 1e-3),
        "unit": "matrix_elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform [-1,1) vector; exact Hamiltonian)",
        "config": {"workload": "heisenberg_chain_L%d_periodic_J1_Sz0_one_vector_hv" % args.length,
                   "dimension": dim, "matrix_elements_per_hv": nnz, "k": 1, "index_lookup": "rank",
                   "device": name, "compute_units": ncu},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / (HBM_PEAK / 1e9),
                     "traffic": (traffic or {}).get("bytes_per_launch"),
                     "kernel": "spin_rank_kernel<1, UNIFORM>", "algorithmic_bytes_per_launch": alg_bytes},
        # the algorithmic model reads x once per matrix element; L2 and MALL serve repeated reads, so
        # the measured HBM bytes (rocprofv3 FETCH_SIZE + WRITE_SIZE) are fewer: their rate is the
        # device's real HBM load, beside `achieved`
        "measured_traffic": traffic and {"bytes_per_launch": traffic["bytes_per_launch"],
                                         "GBps": traffic["bytes_per_launch"] / (ms * 1e-3) / 1e9,
                                         "frac_of_peak": traffic["bytes_per_launch"] / (ms * 1e-3) / HBM_PEAK,
                                         "counted_at_block_size": traffic.get("block_size")},
        "copy_calibration": calib,
        # three readings of the same kernel time: the algorithmic model above (x read once per
        # matrix element, which the caches partly serve), the compulsory bytes (every 32-bit state
        # word, x and y once each: what no implementation can avoid) and the HBM bytes the counters
        # measured (rocprofv3 FETCH_SIZE + WRITE_SIZE)
        "roofline_readings": {
            "algorithmic_frac": achieved / (HBM_PEAK / 1e9),
            "compulsory_bytes_per_launch": dim * (4 + 8 + 8),
            "compulsory_frac": dim * (4 + 8 + 8) / (ms * 1e-3) / HBM_PEAK,
            "counter_frac": traffic and traffic["bytes_per_launch"] / (ms * 1e-3) / HBM_PEAK,
            "counter_over_compulsory": traffic and traffic["bytes_per_launch"] / (dim * (4 + 8 + 8)),
        },
    }
    progress("headline product: %.3f ms" % ms)
    # the LUT-seeded binary search of Spin1_2Chain::FindStateIndex on the same product, beside the
    # combinatorial rank the headline uses
    bs = spin_hv_bench(capi, arch, stream, args.length, min(args.steps, 20), min(args.warmup, 2),
                       index_mode=capi.INDEX_BSEARCH)
    line["index_lookup"] = {"rank_ms": ms, "binary_search_ms": bs["ms_per_step"],
                            "binary_search_elements_per_s": nnz / (bs["ms_per_step"] * 1e-3),
                            "kernel": "spin_bsearch_kernel<1>"}
    progress("binary-search product: %.3f ms" % bs["ms_per_step"])
    if not args.skip_lanczos:
        line["lanczos"] = lanczos_runs(capi, arch, progress)
        n12_legs = (("plain_sequential_k1", 1, False), ("batched_k1", 1, True), ("batched_block_k8", 8, True))
        want = [leg for leg in n12_legs if args.n12 == "all" or leg[0] in args.n12.split(",")]
        if want:
            line["lanczos"]["sectors_n12_2s33"] = sector_runs(capi, arch, progress, n=12, lz=33, nev=10,
                                                              legs=tuple(want), dump=args.n12_dump)
        else:
            recorded = load_traffic(N12_RECORD)
            if recorded:
                line["lanczos"]["sectors_n12_2s33"] = dict(recorded, recorded_not_timed_in_this_run=True,
                                                           source="profiles/" + N12_RECORD)
        if not args.skip_lattice:
            line["lanczos"]["lattice_6x5_n10"] = lattice_runs(capi, arch, progress)
        if not args.skip_n10:
            line["lanczos"]["sectors_n10_2s27"] = sector_runs(capi, arch, progress)
            line["lanczos"]["sectors_autotune_n10_2s27"] = sectors_autotune_leg(capi, arch)
            progress("sector autotuner: %.1f s" % line["lanczos"]["sectors_autotune_n10_2s27"]["seconds"])
        line["reorthogonalisation"] = reorth_stream_leg(capi, arch, stream)
        progress("reorthogonalisation passes: dots %.2f ms, axpy %.2f ms" % (line["reorthogonalisation"]["dots"]["ms"],
                                                                            line["reorthogonalisation"]["axpy"]["ms"]))
        line["adjoint"] = adjoint_leg(capi, arch, stream)
        progress("adjoint: %.3f ms" % line["adjoint"]["adjoint_ms"])
    if not args.skip_batched:
        line["batched_roofline"] = batched_hv_roofline(capi, arch, stream)
        progress("batched product: %.1f ms" % line["batched_roofline"]["ms"])
    if not args.skip_cpu:
        line["cpu_baseline"] = cpu_baseline(args.length)
        progress("cpu baseline done")
    arch.sync()
    print(json.dumps(line))
    sys.stdout.flush()
    arch.close()


if __name__ == "__main__":
    main()
