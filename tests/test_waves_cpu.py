"""CPU checks of the wave planner of the batched sector solves (RunSectorWaves in the host layer),
against a scripted device: waves hold the largest sectors first within the free memory, sectors
whose Krylov space outgrows the storage go to a later wave with more room, and a sector that
cannot fit alone is refused."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def waves(tmp_path_factory):
    lib = os.path.join(ROOT, "diagham_amd", "libdiagham_host.so")
    if not os.path.exists(lib):
        pytest.skip("build the native libraries first")
    exe = str(tmp_path_factory.mktemp("waves") / "waves_selftest")
    libdir = os.path.join(ROOT, "diagham_amd")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fopenmp", "-I", os.path.join(ROOT, "diagham_amd", "host"),
                    os.path.join(ROOT, "tests", "waves_selftest.cc"), "-o", exe, "-L", libdir, "-ldiagham_host",
                    "-Wl,-rpath," + libdir], check=True)
    return exe


def run(exe, tmp_path, k, nev, free, dims, need, kr=None, max_iter=3000):
    lines = ["%d %d %d %d %d %d" % (k, kr or k, nev, max_iter, free, len(dims))]
    lines += ["%d %d" % (d, n) for d, n in zip(dims, need)]
    p = tmp_path / "waves.txt"
    p.write_text("\n".join(lines) + "\n")
    out = subprocess.run([exe, str(p)], stdout=subprocess.PIPE, text=True, check=True).stdout.split("\n")
    if out[0].startswith("error"):
        return out[0], None
    rows = np.array([[int(v) for v in line.split()] for line in out[1:] if line.strip()])
    return int(out[0].split()[1]), rows


def test_all_in_one_wave_when_they_fit(waves, tmp_path):
    dims = [1000, 800, 600, 400]
    w, r = run(waves, tmp_path, 1, 10, 8 * 3000 * 400 * 2, dims, [100, 90, 80, 70])
    assert w == 1 and np.all(r[:, 0] == 1) and np.all(r[:, 1] == 1)


def test_underestimated_sectors_run_again_with_more_room(waves, tmp_path):
    """sectors needing far more blocks than expected stop deferred and converge in a later wave"""
    dims = [1000] * 4 + [100] * 8
    need = [1000] * 4 + [40] * 8          # the expectation for nev = 2, k = 1 is 100 blocks
    free = 8 * (1100 * 1000 + 1000)       # one large sector at a time holds 1000 blocks
    w, r = run(waves, tmp_path, 1, 2, free, dims, need)
    assert np.all(r[:, 0] == 1), r
    assert np.all(r[:4, 2] == 1000) and np.any(r[:4, 1] > 1)
    assert w >= 5


def test_sector_larger_than_the_device_is_refused(waves, tmp_path):
    w, r = run(waves, tmp_path, 1, 10, 8 * 50 * 1000, [1000], [400])
    assert isinstance(w, str) and "more Krylov storage" in w, w


@pytest.mark.parametrize("k", [1, 8])
def test_waves_follow_the_used_blocks(waves, tmp_path, k):
    """133 sectors of decreasing size (the N=12, 2S=33 profile scaled down): every sector converges,
    each at most twice, in few waves"""
    lz = np.arange(133)
    dims = np.maximum(1, (7764392 * np.exp(-(lz / 45.0) ** 2)).astype(int) // 100)
    need = np.minimum(np.maximum(3, (60 + 200 // k) * np.ones(133, int)), dims // k + 1)
    free = 8 * k * 300 * int(dims.sum()) // 4
    w, r = run(waves, tmp_path, k, 10, free, dims, need)
    assert np.all(r[:, 0] == 1)
    assert np.all(r[:, 1] <= 2)
    assert w <= 12


def test_growth_within_the_room_needs_no_restart(waves, tmp_path):
    """Sectors that need 1.3 times the 500 blocks expected for 10 eigenvalues converge in their first
    wave: a wave is admitted with room for 1.5 times its expected storage, so it holds six of the nine
    sectors rather than all nine at the expectation (which would stop all nine at 499 blocks)."""
    dims = [2000] * 9
    need = [650] * 9
    free = 8 * (752 * 12000 + 12000)
    w, r = run(waves, tmp_path, 1, 10, free, dims, need)
    assert np.all(r[:, 0] == 1) and np.all(r[:, 1] == 1), r
    assert w == 2
