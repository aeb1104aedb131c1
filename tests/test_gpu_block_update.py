"""The fused basis update of the two-round Gram-Schmidt (dh_block_update, dh_sectors_block_update)
against numpy on the same host data: W M + sum_b V_b C_b, V^T W and W^T W of the result."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run_single(capi, arch, V, W, C, M, want_dots, k):
    nb, dim = V.shape[0], W.shape[0]
    dV = capi.DeviceBuffer.from_host(arch, V.reshape(-1))
    dW = capi.DeviceBuffer.from_host(arch, W.reshape(-1))
    dots = np.zeros(nb * k * k) if want_dots else None
    gram = np.zeros(k * k)
    rc = capi.gpu().dh_block_update(arch.ctx, dV.ptr, nb, dW.ptr, dim, k, capi.dptr(C.reshape(-1).copy()),
                                    capi.dptr(M.reshape(-1).copy()) if M is not None else None,
                                    capi.dptr(dots) if want_dots else None, capi.dptr(gram))
    out = dW.download().reshape(dim, k)
    dV.free()
    dW.free()
    return rc, out, dots, gram


@pytest.mark.parametrize("k", [1, 2])
@pytest.mark.parametrize("nb", [1, 7, 64, 250])
@pytest.mark.parametrize("with_m", [False, True])
def test_block_update_matches_numpy(arch, k, nb, with_m):
    from diagham_amd import capi
    rng = np.random.default_rng(nb * 10 + k + 3 * with_m)
    dim = 20011
    V = rng.standard_normal((nb, dim, k))
    W = rng.standard_normal((dim, k))
    C = rng.standard_normal((nb, k, k)) * 0.1
    M = rng.standard_normal((k, k)) if with_m else None
    rc, out, dots, gram = run_single(capi, arch, V, W, C, M, True, k)
    assert rc == 0
    ref = (W @ M if with_m else W) + np.einsum("bdp,bpq->dq", V, C)
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()
    rd = np.einsum("bdp,dq->bpq", V, ref).reshape(-1)
    assert np.abs(dots - rd).max() <= 1e-12 * np.abs(rd).max() * np.sqrt(dim)
    rg = (ref.T @ ref).reshape(-1)
    assert np.abs(gram - rg).max() <= 1e-12 * np.abs(rg).max()


def test_block_update_without_dots_and_too_many_blocks(arch):
    from diagham_amd import capi
    rng = np.random.default_rng(5)
    dim, k = 3001, 1
    V = rng.standard_normal((40, dim, k))
    W = rng.standard_normal((dim, k))
    C = rng.standard_normal((40, k, k))
    rc, out, dots, gram = run_single(capi, arch, V, W, C, None, False, k)
    assert rc == 0 and dots is None
    ref = W + np.einsum("bdp,bpq->dq", V, C)
    assert abs(gram[0] - float(ref[:, 0] @ ref[:, 0])) <= 1e-12 * float(ref[:, 0] @ ref[:, 0])
    # a basis too long for the staged tile is refused (the caller then uses the separate kernels)
    Vb = rng.standard_normal((700, 64, k))
    rc, _, _, _ = run_single(capi, arch, Vb, rng.standard_normal((64, k)), rng.standard_normal((700, k, k)), None,
                             True, k)
    assert rc == capi.DH_E_RANGE


@pytest.mark.parametrize("k", [1, 2])
def test_sectors_block_update_matches_numpy(arch, k):
    from diagham_amd import capi
    lz = 15
    lzs = capi.sphere_sectors(6, lz)
    S = capi.SphereSectorSet(arch, 6, lz, lzs, capi.coulomb_pseudopotentials(lz, True))
    total, off = S.total, S.offsets
    active = np.zeros(len(lzs), np.uint8)
    active[::2] = 1
    act = np.flatnonzero(active)
    nb = 9
    rng = np.random.default_rng(k)
    V = rng.standard_normal((nb, total, k))
    W = rng.standard_normal((total, k))
    C = rng.standard_normal((len(act), nb, k, k)) * 0.2
    M = rng.standard_normal((len(act), k, k))
    dV = capi.DeviceBuffer.from_host(arch, V.reshape(-1))
    dW = capi.DeviceBuffer.from_host(arch, W.reshape(-1))
    dots = np.zeros(len(act) * nb * k * k)
    gram = np.zeros(len(act) * k * k)
    rc = capi.gpu().dh_sectors_block_update(S.h, dV.ptr, nb, dW.ptr, k, bytes(active), capi.dptr(C.reshape(-1).copy()),
                                            capi.dptr(M.reshape(-1).copy()), capi.dptr(dots), capi.dptr(gram))
    assert rc == 0
    out = dW.download().reshape(total, k)
    dots = dots.reshape(len(act), nb, k, k)
    gram = gram.reshape(len(act), k, k)
    for s in range(len(lzs)):
        a, b = off[s], off[s + 1]
        if not active[s]:
            assert np.array_equal(out[a:b], W[a:b])
            continue
        j = int(np.searchsorted(act, s))
        ref = W[a:b] @ M[j] + np.einsum("bdp,bpq->dq", V[:, a:b], C[j])
        assert np.abs(out[a:b] - ref).max() <= 1e-12 * np.abs(ref).max()
        assert np.allclose(dots[j], np.einsum("bdp,dq->bpq", V[:, a:b], ref), rtol=0, atol=1e-11 * np.abs(ref).max())
        assert np.allclose(gram[j], ref.T @ ref, rtol=1e-12, atol=1e-12)
    dV.free()
    dW.free()
    S.close()
