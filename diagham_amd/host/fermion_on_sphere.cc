// GPU architecture, device vectors and the Hilbert-space wrappers.
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "diagham_host.hpp"

namespace dhh {

GPUArchitecture::GPUArchitecture(int device) {
  int rc = dh_create(&ctx, device);
  if (rc != DH_OK) throw std::runtime_error("dh_create failed (" + std::to_string(rc) + "): no usable GPU");
}

GPUArchitecture::~GPUArchitecture() {
  if (ctx) dh_destroy(ctx);
}

void GPUArchitecture::Check(int rc, const char* what) const {
  if (rc != DH_OK) throw std::runtime_error(std::string(what) + ": " + dh_last_error(ctx));
}

double* GPUArchitecture::AllocateMemory(int64_t n) {
  void* p = nullptr;
  Check(dh_malloc(ctx, &p, sizeof(double) * (size_t)(n > 0 ? n : 1)), "AllocateMemory");
  return (double*)p;
}

void GPUArchitecture::FreeMemory(double* p) {
  if (p) dh_free(ctx, p);
}

DeviceVector::DeviceVector(GPUArchitecture* a, int64_t n, int kk) : arch(a), dim(n), k(kk) {
  d = arch->AllocateMemory(dim * k);
  arch->Check(dh_memset(arch->Context(), d, 0, sizeof(double) * dim * k), "DeviceVector clear");
}

DeviceVector::~DeviceVector() {
  if (arch && d && owns) arch->FreeMemory(d);
}

DeviceVector::DeviceVector(DeviceVector&& o) noexcept : arch(o.arch), d(o.d), dim(o.dim), k(o.k), owns(o.owns) {
  o.d = nullptr;
}

DeviceVector& DeviceVector::operator=(DeviceVector&& o) noexcept {
  if (this != &o) {
    if (arch && d && owns) arch->FreeMemory(d);
    arch = o.arch;
    d = o.d;
    dim = o.dim;
    k = o.k;
    owns = o.owns;
    o.d = nullptr;
  }
  return *this;
}

DeviceVector DeviceVector::View(GPUArchitecture* a, double* p, int64_t n, int kk) {
  DeviceVector v;
  v.arch = a;
  v.d = p;
  v.dim = n;
  v.k = kk;
  v.owns = false;
  return v;
}

AbstractHilbertSpace::~AbstractHilbertSpace() {
  if (basis) dh_basis_destroy(basis);
}

FermionOnSphere::FermionOnSphere(GPUArchitecture* arch, int nbrFermions, int totalLz, int lzMax)
    : NbrFermions(nbrFermions), TotalLz(totalLz), LzMax(lzMax) {
  arch->Check(dh_fermion_sphere_create(arch->Context(), &basis, nbrFermions, totalLz, lzMax), "FermionOnSphere");
}

Spin1_2Chain::Spin1_2Chain(GPUArchitecture* arch, int chainLength, int sz) : ChainLength(chainLength), Sz(sz) {
  arch->Check(dh_spin_chain_create(arch->Context(), &basis, chainLength, sz), "Spin1_2Chain");
}

AbstractHamiltonian::~AbstractHamiltonian() {
  if (ham) dh_hamiltonian_destroy(ham);
}

void AbstractHamiltonian::ShiftHamiltonian(double shift) {
  Shift = shift;
  arch->Check(dh_hamiltonian_set_shift(ham, shift), "ShiftHamiltonian");
}

DeviceVector& AbstractHamiltonian::LowLevelAddMultiply(DeviceVector& src, DeviceVector& dst, int64_t first,
                                                       int64_t nbr) {
  arch->Check(dh_hv(ham, src.Data(), dst.Data(), src.BlockSize(), 1, first, nbr), "LowLevelAddMultiply");
  return dst;
}

DeviceVector& AbstractHamiltonian::LowLevelMultiply(DeviceVector& src, DeviceVector& dst, int64_t first, int64_t nbr) {
  arch->Check(dh_hv(ham, src.Data(), dst.Data(), src.BlockSize(), 0, first, nbr), "LowLevelMultiply");
  return dst;
}

DeviceVector& AbstractHamiltonian::LowLevelMultipleAddMultiply(DeviceVector& src, DeviceVector& dst) {
  arch->Check(dh_hv(ham, src.Data(), dst.Data(), src.BlockSize(), 1, 0, -1), "LowLevelMultipleAddMultiply");
  return dst;
}

DeviceVector& AbstractHamiltonian::LowLevelMultipleMultiply(DeviceVector& src, DeviceVector& dst) {
  arch->Check(dh_hv(ham, src.Data(), dst.Data(), src.BlockSize(), 0, 0, -1), "LowLevelMultipleMultiply");
  return dst;
}

static DeviceVector* MultipleProduct(AbstractHamiltonian* h, GPUArchitecture* arch, DeviceVector* src,
                                     DeviceVector* dst, int n, int add, int64_t first, int64_t nbr, const char* what) {
  std::vector<const double*> x(n > 0 ? n : 1);
  std::vector<double*> y(n > 0 ? n : 1);
  for (int v = 0; v < n; ++v) {
    if (src[v].BlockSize() != 1 || dst[v].BlockSize() != 1)
      throw std::invalid_argument(std::string(what) + ": the vectors must be single vectors");
    x[v] = src[v].Data();
    y[v] = dst[v].Data();
  }
  arch->Check(dh_hv_vectors(h->Handle(), x.data(), y.data(), n, add, first, nbr), what);
  return dst;
}

DeviceVector* AbstractHamiltonian::LowLevelMultipleAddMultiply(DeviceVector* src, DeviceVector* dst, int n,
                                                               int64_t first, int64_t nbr) {
  return MultipleProduct(this, arch, src, dst, n, 1, first, nbr, "LowLevelMultipleAddMultiply");
}

DeviceVector* AbstractHamiltonian::LowLevelMultipleMultiply(DeviceVector* src, DeviceVector* dst, int n,
                                                            int64_t first, int64_t nbr) {
  return MultipleProduct(this, arch, src, dst, n, 0, first, nbr, "LowLevelMultipleMultiply");
}

MultipleVectorHamiltonianMultiplyOperation::MultipleVectorHamiltonianMultiplyOperation(AbstractHamiltonian* h,
                                                                                       DeviceVector* s,
                                                                                       DeviceVector* d, int n)
    : Hamiltonian(h), Sources(s), Destinations(d), NbrVectors(n) {}

void MultipleVectorHamiltonianMultiplyOperation::SetIndicesRange(int64_t first, int64_t nbr) {
  FirstComponent = first;
  NbrComponent = nbr;
}

bool MultipleVectorHamiltonianMultiplyOperation::ApplyOperation(GPUArchitecture*) {
  Hamiltonian->LowLevelMultipleMultiply(Sources, Destinations, NbrVectors, FirstComponent, NbrComponent);
  return true;
}

VectorHamiltonianMultiplyOperation::VectorHamiltonianMultiplyOperation(AbstractHamiltonian* h, DeviceVector* s,
                                                                       DeviceVector* d)
    : Hamiltonian(h), Source(s), Destination(d) {}

bool VectorHamiltonianMultiplyOperation::ApplyOperation(GPUArchitecture*) {
  Hamiltonian->LowLevelMultiply(*Source, *Destination, 0, Source->GetVectorDimension());
  return true;
}

}  // namespace dhh
