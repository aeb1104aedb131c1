"""CPU checks of the tridiagonal routines of the one-vector Lanczos in the host layer: the inverse
iteration that gives the residual test its Ritz vectors' last components (TridiagonalEigenvector),
against numpy's dense eigenvectors on random and Lanczos-shaped tridiagonal matrices, including a
close pair and a matrix whose elimination needs row exchanges; and the Sturm-count bisection of the
convergence test's lowest eigenvalues (TridiagonalLowest), against numpy's dense eigenvalues, with
the previous step's eigenvalues as bracket hints changing no bit of the results."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def tridiag(tmp_path_factory):
    lib = os.path.join(ROOT, "diagham_amd", "libdiagham_host.so")
    if not os.path.exists(lib):
        pytest.skip("build the native libraries first")
    exe = str(tmp_path_factory.mktemp("tridiag") / "tridiag_selftest")
    libdir = os.path.join(ROOT, "diagham_amd")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fopenmp", "-I", os.path.join(ROOT, "diagham_amd", "host"),
                    os.path.join(ROOT, "tests", "tridiag_selftest.cc"), "-o", exe, "-L", libdir, "-ldiagham_host",
                    "-Wl,-rpath," + libdir], check=True)
    return exe


def eigenvector(exe, tmp_path, d, e, theta):
    p = tmp_path / "t.txt"
    p.write_text("%d %.17g\n" % (len(d), theta) + " ".join("%.17g" % v for v in d) + "\n" +
                 " ".join("%.17g" % v for v in e) + "\n")
    out = subprocess.run([exe, str(p)], stdout=subprocess.PIPE, text=True, check=True).stdout.split()
    return np.array([float(v) for v in out])


def check(exe, tmp_path, d, e, which):
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    w, V = np.linalg.eigh(T)
    for i in which:
        z = eigenvector(exe, tmp_path, d, e, w[i])
        assert abs(np.linalg.norm(z) - 1.0) < 1e-12
        # the residual of the computed vector, and its agreement with the dense eigenvector
        # (up to sign) where the eigenvalue is separated from the others
        assert np.linalg.norm(T @ z - w[i] * z) < 1e-12 * np.abs(w).max()
        gap = np.min(np.abs(np.delete(w, i) - w[i]))
        if gap > 1e-6 * np.abs(w).max():
            s = np.sign(z @ V[:, i])
            assert np.abs(z - s * V[:, i]).max() < 1e-10 * np.abs(w).max() / gap
            assert abs(abs(z[-1]) - abs(V[-1, i])) < 1e-10 * np.abs(w).max() / gap


def test_random_tridiagonal(tridiag, tmp_path):
    rng = np.random.default_rng(3)
    d, e = rng.standard_normal(200), rng.standard_normal(199)
    check(tridiag, tmp_path, d, e, [0, 1, 2, 100, 199])


def test_lanczos_shaped_matrix(tridiag, tmp_path):
    """the projection of a wide spectrum: the lowest eigenvectors' last components are tiny, and
    they are what the residual test reads"""
    rng = np.random.default_rng(4)
    A = np.diag(np.concatenate([[0.0, 0.3, 0.31], rng.uniform(1.0, 800.0, 597)]))
    v = rng.standard_normal(600)
    v /= np.linalg.norm(v)
    Q = [v]
    d, e = [], []
    for j in range(120):
        w = A @ Q[-1]
        d.append(Q[-1] @ w)
        for q in Q:
            w -= (q @ w) * q
        for q in Q:
            w -= (q @ w) * q
        b = np.linalg.norm(w)
        if j < 119:
            e.append(b)
            Q.append(w / b)
    check(tridiag, tmp_path, np.array(d), np.array(e), [0, 1, 2])


def test_zero_diagonal_needs_row_exchanges(tridiag, tmp_path):
    d = np.zeros(9)
    e = np.arange(1.0, 9.0)
    check(tridiag, tmp_path, d, e, range(9))


def lowest(exe, tmp_path, d, e, count, hint=()):
    p = tmp_path / "l.txt"
    p.write_text("%d %d %d\n" % (len(d), count, len(hint)) + " ".join("%.17g" % v for v in d) + "\n" +
                 " ".join("%.17g" % v for v in e) + "\n" + " ".join("%.17g" % v for v in hint) + "\n")
    return subprocess.run([exe, str(p), "lowest"], stdout=subprocess.PIPE, text=True, check=True).stdout.split()


def lanczos_tridiagonal(rng, n=400, m=160):
    """the tridiagonal matrix of m full-reorthogonalised Lanczos steps on a spectrum with a dense,
    nearly degenerate low end (as in a many-body sector)"""
    ev = np.sort(np.concatenate([5 + np.cumsum(rng.exponential(0.02, 30)), rng.uniform(6, 30, n - 30)]))
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (Q * ev) @ Q.T
    V = np.zeros((n, m + 1))
    d, e = np.zeros(m), np.zeros(m)
    v = rng.standard_normal(n)
    V[:, 0] = v / np.linalg.norm(v)
    for j in range(m):
        w = A @ V[:, j]
        for _ in range(2):
            w -= V[:, :j + 1] @ (V[:, :j + 1].T @ w)
        d[j] = V[:, j] @ (A @ V[:, j])
        e[j] = np.linalg.norm(w)
        V[:, j + 1] = w / e[j]
    return d, e[:m - 1]


@pytest.mark.parametrize("kind", ["random", "clustered", "lanczos"])
def test_lowest_eigenvalues_match_dense(tridiag, tmp_path, kind):
    rng = np.random.default_rng(9)
    if kind == "random":
        d, e = rng.standard_normal(150), rng.standard_normal(149)
    elif kind == "clustered":
        d, e = np.ones(120), 1e-9 * rng.standard_normal(119)
    else:
        d, e = lanczos_tridiagonal(rng)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    w = np.linalg.eigvalsh(T)
    got = np.array([float(v) for v in lowest(tridiag, tmp_path, d, e, 20)])
    # both are backward stable: O(n eps |T|) apart at most, far closer in practice
    assert np.abs(got - w[:20]).max() <= 32 * np.finfo(float).eps * np.abs(w).max()
    assert np.all(np.diff(got) >= 0)


def test_lowest_hints_change_no_bit(tridiag, tmp_path):
    """Along a growing Lanczos matrix, the previous step's values as hints (Cauchy interlacing
    bounds) give the same doubles as the bisection from the Gershgorin bracket; so do hints that
    bound nothing (random values, a hint list longer or shorter than the count)."""
    rng = np.random.default_rng(10)
    d, e = lanczos_tridiagonal(rng, m=120)
    prev = ()
    for m in range(12, 121, 9):
        plain = lowest(tridiag, tmp_path, d[:m], e[:m - 1], 11)
        hinted = lowest(tridiag, tmp_path, d[:m], e[:m - 1], 11, prev)
        assert hinted == plain, m
        prev = [float(v) for v in hinted]
    junk = rng.uniform(-50, 50, 15)
    assert lowest(tridiag, tmp_path, d, e, 11, junk) == lowest(tridiag, tmp_path, d, e, 11)
    assert lowest(tridiag, tmp_path, d, e, 11, prev[:4]) == lowest(tridiag, tmp_path, d, e, 11)
