"""Disk storage of the Lanczos run (FullReorthogonalizedLanczosAlgorithmWithDiskStorage,
--disk / --resume of GenericRealMainTask.cc:120-121): Krylov vectors as DiagHam vector files and
the recurrence state in lanczos.dat, and a resumed run that repeats an uninterrupted one."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def read_vec(path):
    raw = np.fromfile(path, dtype=np.uint8)
    d = int(np.frombuffer(raw[:4], "<i4")[0])
    assert d >= 0
    v = np.frombuffer(raw[4:], "<f8")
    assert v.size == d
    return v


@pytest.mark.parametrize("k", [1, 4])
def test_resumed_run_repeats_the_uninterrupted_one(arch, tmp_path, golden, k):
    from diagham_amd import capi
    pp = O.round14(O.coulomb_pseudopotentials(21))
    p = capi.Problem.sphere(arch, 8, 0, 21, pp)
    full = p.lanczos(10, block_size=k, max_iter=2000)
    assert full["converged"]
    # first leg: stopped by max_iter well before convergence, every vector on disk
    d = tmp_path / ("k%d" % k)
    d.mkdir()
    first = p.lanczos(10, block_size=k, max_iter=12 * k, disk=str(d), allow_unconverged=True)
    assert not first["converged"]
    nvec = (first["blocks"] + 1) * k
    files = sorted(x.name for x in d.iterdir())
    assert "lanczos.dat" in files and "vector.%d" % (nvec - 1) in files and "vector.%d" % nvec not in files
    # the stored Krylov vectors are orthonormal vectors of the space
    V = np.stack([read_vec(d / ("vector.%d" % i)) for i in range(nvec)])
    assert V.shape == (nvec, p.dim)
    assert np.abs(V @ V.T - np.eye(nvec)).max() < 1e-12
    # second leg: resume with the full iteration budget
    second = p.lanczos(10, block_size=k, max_iter=2000, disk=str(d), resume=True)
    assert second["converged"]
    assert np.array_equal(second["eigenvalues"], full["eigenvalues"])
    assert second["blocks"] == full["blocks"]
    ref = np.array(golden["fermions_sphere_n8_2s21_lz0"]["eigenvalues"])[:10]
    assert np.abs(second["eigenvalues"][:ref.size] - ref).max() < 1e-10
    p.close()


def test_resume_refuses_another_space(arch, tmp_path):
    from diagham_amd import capi
    pp = O.round14(O.coulomb_pseudopotentials(21))
    p = capi.Problem.sphere(arch, 8, 0, 21, pp)
    p.lanczos(4, max_iter=10, disk=str(tmp_path), allow_unconverged=True)
    q = capi.Problem.sphere(arch, 8, 2, 21, pp)
    with pytest.raises(capi.DHError, match="does not match dimension of stored vectors"):
        q.lanczos(4, max_iter=100, disk=str(tmp_path), resume=True)
    with pytest.raises(capi.DHError, match="cannot open"):
        p.lanczos(4, max_iter=100, disk=str(tmp_path / "missing"), resume=True)
    q.close()
    p.close()
