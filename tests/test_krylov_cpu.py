"""CPU checks of the host layer's block Krylov recurrence (rank deficiency, small and degenerate
spaces, complex Hermitian blocks): the recurrence is compiled against the host library with a
host-memory implementation of its block operations and compared with numpy's dense eigensolver."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def selftest(tmp_path_factory):
    lib = os.path.join(ROOT, "diagham_amd", "libdiagham_host.so")
    if not os.path.exists(lib):
        pytest.skip("build the native libraries first")
    exe = str(tmp_path_factory.mktemp("krylov") / "krylov_selftest")
    libdir = os.path.join(ROOT, "diagham_amd")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fopenmp", "-I", os.path.join(ROOT, "diagham_amd", "host"),
                    os.path.join(ROOT, "tests", "krylov_selftest.cc"), "-o", exe, "-L", libdir, "-ldiagham_host",
                    "-Wl,-rpath," + libdir], check=True)
    return exe


def run(exe, tmp_path, mats, k, nev, precision=1e-14, max_iter=4000, seed=11, mode=None, raw=False):
    cplx = any(np.iscomplexobj(m) for m in mats)
    lines = ["%d %d %d %d %.17g %d %d" % (cplx, k, max_iter, nev, precision, seed, len(mats)),
             " ".join(str(m.shape[0]) for m in mats)]
    for m in mats:
        if cplx:
            m = np.asarray(m, complex)
            flat = np.empty(2 * m.size)
            flat[0::2] = m.real.ravel()
            flat[1::2] = m.imag.ravel()
        else:
            flat = m.ravel()
        lines.append(" ".join(repr(float(v)) for v in flat))
    path = tmp_path / "problem.txt"
    path.write_text("\n".join(lines) + "\n")
    out = subprocess.run([exe, str(path)] + ([mode] if mode else []), stdout=subprocess.PIPE,
                         stderr=subprocess.PIPE, text=True)
    assert out.returncode == 0, out.stderr
    if raw:
        return out
    res = []
    for line in out.stdout.strip().splitlines():
        v = line.split()
        res.append((int(v[0]), int(v[1]), np.array([float(x) for x in v[2:]])))
    return res


def random_hermitian(rng, d, cplx=False):
    a = rng.standard_normal((d, d))
    if cplx:
        a = a + 1j * rng.standard_normal((d, d))
    return (a + a.conj().T) / 2


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("k", [1, 4, 8, 16])
def test_sectors_of_every_size(selftest, tmp_path, k, cplx):
    """Sectors from 1 state to well above the block size converge to the dense spectrum, also when
    the Krylov space fills a sector in the middle of a block."""
    rng = np.random.default_rng(k + 7 * cplx)
    dims = [1, 2, 3, 5, 9, 17, 33, 70, 130]
    mats = [random_hermitian(rng, d, cplx) for d in dims]
    nev = 5
    res = run(selftest, tmp_path, mats, k, nev)
    for m, (conv, prods, ev) in zip(mats, res):
        ref = np.linalg.eigvalsh(m)[:nev]
        assert conv == 1, m.shape
        assert ev.size == ref.size
        assert np.max(np.abs(ev - ref)) < 1e-10 * max(1.0, np.max(np.abs(ref))), (m.shape, ev, ref)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("k", [1, 2, 8])
def test_growing_krylov_storage_repeats_the_run(selftest, tmp_path, k, cplx):
    """Storage that starts with two Krylov blocks and grows on demand by further chunks (fresh
    buffers filled with poison) gives the eigenvalues of storage for every block from the start,
    through convergence, exhaustion and the random restarts of a degenerate sector.  The products
    over the basis then run chunk by chunk, which changes only the grouping of their sums: the
    eigenvalues agree to rounding and the runs converge alike."""
    rng = np.random.default_rng(300 + k + cplx)
    mats = [random_hermitian(rng, d, cplx) for d in (3, 17, 70, 130)]
    lam = np.concatenate([np.full(9, -3.0), rng.uniform(-2, 5, 31)])
    q, _ = np.linalg.qr(rng.standard_normal((40, 40)) + (1j * rng.standard_normal((40, 40)) if cplx else 0))
    mats.append((q * lam) @ q.conj().T)
    for nev, precision in ((5, 1e-14), (40, 0.0)):
        fixed = run(selftest, tmp_path, mats, k, nev, precision=precision, raw=True)
        grown = run(selftest, tmp_path, mats, k, nev, precision=precision, mode="grow", raw=True)
        rows_f = [line.split() for line in fixed.stdout.strip().splitlines()]
        rows_g = [line.split() for line in grown.stdout.strip().splitlines()]
        assert len(rows_f) == len(rows_g)
        for f, g in zip(rows_f, rows_g):
            assert f[0] == g[0]  # converged alike
            assert abs(int(f[1]) - int(g[1])) <= 2 * k  # a convergence test may fall one block apart
            if f[0] == "1":  # a sector stopped unconverged at its block cap with random restarts
                ef, eg = np.array(f[2:], float), np.array(g[2:], float)  # follows its rounding
                assert ef.size == eg.size and np.max(np.abs(ef - eg), initial=0.0) < 1e-11
        growths, chunks = int(grown.stderr.split()[1]), int(grown.stderr.split()[5])
        assert growths >= 2 and chunks == growths + 1, grown.stderr


@pytest.mark.parametrize("cplx", [False, True])
def test_split_step_refuses_misuse(selftest, tmp_path, cplx):
    """FinishStep needs the product of its own step, launched for every segment it
    orthogonalises; a refused call changes nothing, and a proper one makes exactly one step."""
    rng = np.random.default_rng(5 + cplx)
    mats = [random_hermitian(rng, d, cplx) for d in (12, 30)]
    out = run(selftest, tmp_path, mats, 2, 4, mode="misuse", raw=True)
    assert out.stdout.split() == ["refused", "refused", "accepted", "refused", "steps", "1"], out.stdout


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("k", [2, 4, 8])
def test_run_to_exhaustion_recovers_degenerate_spectrum(selftest, tmp_path, k, cplx):
    """precision 0 runs every sector until its Krylov space is complete: the whole spectrum,
    including an eigenvalue of multiplicity above k (reached only through the restarts that
    replace the columns of an invariant Krylov space), comes out with its multiplicities."""
    rng = np.random.default_rng(100 + k + cplx)
    mats = []
    for d, mult in [(24, 9), (41, 12), (60, 3)]:
        lam = np.concatenate([np.full(mult, -3.0), rng.uniform(-2, 5, d - mult)])
        q, _ = np.linalg.qr(rng.standard_normal((d, d)) + (1j * rng.standard_normal((d, d)) if cplx else 0))
        mats.append((q * lam) @ q.conj().T)
    res = run(selftest, tmp_path, mats, k, nev=60, precision=0.0)
    for m, (conv, prods, ev) in zip(mats, res):
        d = m.shape[0]
        ref = np.linalg.eigvalsh(m)
        assert conv == 1
        assert ev.size == d
        assert np.max(np.abs(ev - ref)) < 1e-10, (d, ev[:14], ref[:14])


@pytest.mark.parametrize("k", [1, 4])
def test_zero_and_scalar_operators(selftest, tmp_path, k):
    """H = 0 and H = c I make every new block vanish after reorthogonalisation: each step is a
    restart, and the spectrum is still exact."""
    mats = [np.zeros((20, 20)), 2.5 * np.eye(37)]
    res = run(selftest, tmp_path, mats, k, nev=6, precision=0.0)
    for m, (conv, prods, ev) in zip(mats, res):
        assert conv == 1
        assert np.max(np.abs(ev - np.linalg.eigvalsh(m)[:ev.size])) < 1e-12
        assert ev.size == 6


@pytest.mark.parametrize("seed", range(6))
def test_random_sector_sets(selftest, tmp_path, seed):
    """Random sector sets (sizes around multiples of k, spectra with clusters) at the default
    precision: the lowest eigenvalues of every sector agree with the dense ones."""
    rng = np.random.default_rng(1000 + seed)
    k = int(rng.choice([1, 2, 4, 8, 16]))
    cplx = bool(seed % 2)
    mats = []
    for _ in range(5):
        d = int(max(1, k * rng.integers(1, 12) + rng.integers(-1, 2)))
        lam = np.sort(rng.uniform(-1, 1, d))
        lam[: min(d, 3)] = lam[0] + 1e-7 * np.arange(min(d, 3))  # a tight cluster at the bottom
        q, _ = np.linalg.qr(rng.standard_normal((d, d)) + (1j * rng.standard_normal((d, d)) if cplx else 0))
        mats.append((q * lam) @ q.conj().T)
    nev = int(rng.integers(1, 8))
    res = run(selftest, tmp_path, mats, k, nev=nev, seed=seed)
    for m, (conv, prods, ev) in zip(mats, res):
        ref = np.linalg.eigvalsh(m)[:nev]
        assert conv == 1
        assert ev.size == ref.size
        assert np.max(np.abs(ev - ref)) < 1e-10, (k, m.shape, ev, ref)


def test_invariant_block_structure(selftest, tmp_path):
    """A block-diagonal operator with a one-dimensional block: early invariant subspaces in the
    middle of the recurrence."""
    rng = np.random.default_rng(5)
    a = random_hermitian(rng, 30)
    m = np.zeros((61, 61))
    m[:30, :30] = a
    m[30:60, 30:60] = a + 0.25
    m[60, 60] = -9.0
    res = run(selftest, tmp_path, [m], 4, nev=8)
    conv, prods, ev = res[0]
    assert conv == 1
    assert np.max(np.abs(ev - np.linalg.eigvalsh(m)[:8])) < 1e-10


def run_resume(exe, tmp_path, mats, k, nev, max_iter, seed=11):
    cplx = any(np.iscomplexobj(m) for m in mats)
    lines = ["%d %d %d %d %.17g %d %d" % (cplx, k, max_iter, nev, 1e-14, seed, len(mats)),
             " ".join(str(m.shape[0]) for m in mats)]
    for m in mats:
        m = np.asarray(m, complex if cplx else float)
        flat = np.stack([m.real.ravel(), m.imag.ravel()], 1).ravel() if cplx else m.ravel()
        lines.append(" ".join(repr(float(v)) for v in flat))
    path = tmp_path / "problem.txt"
    path.write_text("\n".join(lines) + "\n")
    out = subprocess.run([exe, str(path), "resume"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert out.returncode == 0, out.stderr
    return out.stdout.strip().splitlines()


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("k", [1, 4])
def test_checkpoint_resume_repeats_the_run(selftest, tmp_path, k, cplx):
    """SaveState/LoadState (the host half of the disk storage of
    FullReorthogonalizedLanczosAlgorithmWithDiskStorage.cc:472-530): an engine that loads the
    state saved halfway, with only the saved Krylov blocks in fresh storage, ends on Ritz values
    bitwise equal to an uninterrupted run's.  Segments: a generic one, and a degenerate one whose
    Krylov space becomes invariant early, so random restarts (seeded from the saved state) occur
    after the checkpoint."""
    rng = np.random.default_rng(5 + k)
    generic = random_hermitian(rng, 60, cplx)
    q, _ = np.linalg.qr(random_hermitian(rng, 40, cplx))
    degenerate = (q * np.repeat([-1.0, 0.5, 2.0], [10, 20, 10])) @ q.conj().T
    lines = run_resume(selftest, tmp_path, [generic, degenerate], k, nev=3, max_iter=12 * k)
    assert lines[-1].split()[0] == "same", lines
    steps, half = map(int, lines[-1].split()[1:])
    assert steps >= 8 and 0 < half < steps
    # the restarts carried the degenerate segment into its multiplicity-10 lowest eigenspace
    ev = [float(x) for x in lines[1].split()[2::2]]
    assert np.allclose(ev[:3], [-1.0, -1.0, -1.0], atol=1e-10)


@pytest.mark.parametrize("case", ["random", "clustered", "graded", "zero_couplings"])
def test_tridiagonal_lowest_matches_dense(selftest, tmp_path, case):
    """TridiagonalLowest (Sturm bisection used by the one-vector convergence tests) against
    numpy's dense eigvalsh on tridiagonal matrices, including multiple eigenvalues, couplings that
    vanish (a dead Krylov column) and entries spanning many orders of magnitude."""
    rng = np.random.default_rng(["random", "clustered", "graded", "zero_couplings"].index(case))
    n = 237
    d = rng.standard_normal(n) * 10
    e = rng.standard_normal(n - 1)
    if case == "clustered":
        d = np.round(d) / 4
        e *= 1e-9
    elif case == "graded":
        d = d * np.logspace(-6, 3, n)
        e = e * np.logspace(-6, 3, n - 1)
    elif case == "zero_couplings":
        e[::7] = 0.0
        d[-1] = 1e3
    path = tmp_path / "tri.txt"
    count = 11
    path.write_text("%d %d\n%s\n%s\n" % (n, count, " ".join(repr(float(v)) for v in d),
                                         " ".join(repr(float(v)) for v in e)))
    out = subprocess.run([selftest, str(path), "lowest"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert out.returncode == 0, out.stderr
    got = np.array([float(x) for x in out.stdout.split()])
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    ref = np.linalg.eigvalsh(T)[:count]
    scale = np.abs(np.linalg.eigvalsh(T)).max()
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 4e-15 * scale * np.sqrt(n)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("k", [1, 2, 4])
def test_krylov_basis_stays_orthonormal(selftest, tmp_path, k, cplx):
    """After 60 vectors of a clustered spectrum the stored Krylov basis is orthonormal to rounding
    level, for real and complex blocks (the second Gram-Schmidt round of k = 1 reuses the first
    update's dots, divided by the norm: a wrong scale there leaves basis components behind)."""
    rng = np.random.default_rng(31 + k + 5 * cplx)
    d = 220
    lam = np.concatenate([rng.uniform(-1, -0.99, 20), rng.uniform(0, 3, d - 20)])
    q, _ = np.linalg.qr(rng.standard_normal((d, d)) + (1j * rng.standard_normal((d, d)) if cplx else 0))
    m = (q * lam) @ q.conj().T
    lines = [("%d %d %d %d %.17g %d %d" % (cplx, k, 60, 4, 1e-14, 3, 1)), str(d)]
    mm = np.asarray(m, complex if cplx else float)
    flat = np.stack([mm.real.ravel(), mm.imag.ravel()], 1).ravel() if cplx else mm.ravel()
    lines.append(" ".join(repr(float(v)) for v in flat))
    path = tmp_path / "p.txt"
    path.write_text("\n".join(lines) + "\n")
    out = subprocess.run([selftest, str(path), "ortho"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert out.returncode == 0, out.stderr
    assert float(out.stdout.split()[0]) < 1e-12


@pytest.mark.parametrize("k", [1, 4])
def test_exact_zero_mode_converges(selftest, tmp_path, k):
    """A wanted eigenvalue that is exactly zero converges (FullReorthogonalizedLanczosAlgorithm.cc:
    335-338): its Ritz value only moves by rounding noise, which is never below 1e-14 of itself.
    The space is larger than max_iter, so only the convergence test can stop the run."""
    rng = np.random.default_rng(41 + k)
    d, max_iter = 300, 250
    lam = np.concatenate([[0.0, 1.0, 1.5, 2.0], rng.uniform(3.0, 5.0, d - 4)])
    q, _ = np.linalg.qr(rng.standard_normal((d, d)))
    m = (q * lam) @ q.T
    m = (m + m.T) / 2
    nev = 3
    (conv, prods, ev), = run(selftest, tmp_path, [m], k, nev, max_iter=max_iter)
    assert conv == 1 and prods < max_iter, (conv, prods)
    assert abs(ev[0]) < 1e-13
    assert np.max(np.abs(ev - lam[:nev])) < 1e-10
