"""diagham_amd: MI355X-native on-the-fly Hamiltonian-vector product and Lanczos for
DiagHam-style exact diagonalisation.

Native pieces (built in-tree by ``python -m diagham_amd.build``):
  libdhgpu.so         gfx950 HIP kernels behind the C ABI of include/dhgpu.h
  libdiagham_host.so  C++ host layer shaped like DiagHam's classes (include/diagham_host.h)
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
