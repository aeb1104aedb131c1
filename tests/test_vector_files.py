"""DiagHam binary vector files (RealVector::WriteVector src/Vector/RealVector.cc:3669,
ComplexVector::WriteVector src/Vector/ComplexVector.cc:2606): header parsing on the CPU, and a
round trip of a Lanczos eigenvector through device memory on the GPU."""
import ctypes

import numpy as np
import pytest


def write_reference_layout(path, data, complex_=False, large=False):
    with open(path, "wb") as f:
        dim = data.size // (2 if complex_ else 1)
        if large:
            f.write(np.int32(-1).tobytes())
            f.write(np.int64(dim).tobytes())
        else:
            f.write(np.int32(dim).tobytes())
        f.write(np.ascontiguousarray(data, dtype="<f8").tobytes())


@pytest.mark.parametrize("large", [False, True])
def test_vector_file_dimension(tmp_path, large):
    from diagham_amd import capi
    p = tmp_path / "v.vec"
    write_reference_layout(p, np.arange(10.0), large=large)
    d = ctypes.c_int64()
    assert capi.host().dhh_vector_file_dimension(str(p).encode(), 0, ctypes.byref(d)) == 0 and d.value == 10
    # read as complex: 10 doubles hold 5 complex components, not 10
    assert capi.host().dhh_vector_file_dimension(str(p).encode(), 1, ctypes.byref(d)) == capi.DH_E_ARG
    q = tmp_path / "c.vec"
    write_reference_layout(q, np.arange(12.0), complex_=True, large=large)
    assert capi.host().dhh_vector_file_dimension(str(q).encode(), 1, ctypes.byref(d)) == 0 and d.value == 6
    assert capi.host().dhh_vector_file_dimension(str(q).encode(), 0, ctypes.byref(d)) == capi.DH_E_ARG
    assert b"complex" in capi.host().dhh_last_error()


def test_missing_file_is_an_error(tmp_path):
    from diagham_amd import capi
    d = ctypes.c_int64()
    assert capi.host().dhh_vector_file_dimension(str(tmp_path / "none.vec").encode(), 0, ctypes.byref(d)) != 0


@pytest.mark.gpu
def test_eigenvector_round_trip(arch, tmp_path):
    from diagham_amd import capi
    import oracle_lib as O
    lz = 15
    p = capi.Problem.sphere(arch, 6, 0, lz, O.round14(O.coulomb_pseudopotentials(lz)))
    vec = capi.DeviceBuffer(arch, 2 * p.dim)
    p.lanczos(2, block_size=1, d_vectors=vec.ptr, nbr_vectors=2)
    v = vec.download()[:p.dim]
    path = tmp_path / "ground.vec"
    assert capi.host().dhh_write_vector(arch.h, str(path).encode(), vec.ptr, p.dim, 0) == 0
    raw = path.read_bytes()
    assert np.frombuffer(raw[:4], "<i4")[0] == p.dim
    assert np.array_equal(np.frombuffer(raw[4:], "<f8"), v)
    back = capi.DeviceBuffer(arch, p.dim)
    assert capi.host().dhh_read_vector(arch.h, str(path).encode(), back.ptr, p.dim, 0) == 0
    assert np.array_equal(back.download(), v)
    assert capi.host().dhh_read_vector(arch.h, str(path).encode(), back.ptr, p.dim + 1, 0) == capi.DH_E_ARG
    p.close()


@pytest.mark.gpu
@pytest.mark.parametrize("large,complex_", [(False, False), (True, False), (True, True)])
def test_read_either_header(arch, tmp_path, large, complex_):
    """A small vector written with the -1 / int64 header (allowed by the format for any dimension) is
    read from the component after that 12-byte header, not from byte 4."""
    from diagham_amd import capi
    dim = 1001
    data = np.random.default_rng(11).standard_normal(dim * (2 if complex_ else 1))
    path = tmp_path / "v.vec"
    write_reference_layout(path, data, complex_=complex_, large=large)
    back = capi.DeviceBuffer(arch, data.size)
    assert capi.host().dhh_read_vector(arch.h, str(path).encode(), back.ptr, dim, int(complex_)) == 0
    assert np.array_equal(back.download(), data)
    back.free()
