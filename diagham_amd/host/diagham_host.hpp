// C++ host side of the on-the-fly H.v + Lanczos path, shaped like DiagHam's own classes
// (same names, argument meanings and call order) and driving the GPU through the C ABI of
// include/dhgpu.h.  Vectors live in device memory for their whole life.
#pragma once

#include <algorithm>
#include <chrono>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "../../include/dhgpu.h"
#include "block_krylov.hpp"

namespace dhh {

// ------------------------------------------------------------------ math tools
// ClebschGordanCoefficients (src/MathTools/ClebschGordanCoefficients.cc:225): all
// <j1/2 m1/2 ; j2/2 m2/2 | j/2 (m1+m2)/2> for doubled angular momenta.  Evaluated with
// Racah's closed formula in 113-bit floating point, so each value is correctly rounded to
// double in practice (the reference uses the Schulten-Gordon recursion in double).
class ClebschGordanCoefficients {
 public:
  ClebschGordanCoefficients(int j1, int j2);
  // zero outside the allowed range, like GetCoefficient (:150)
  double GetCoefficient(int m1, int m2, int j) const;
  // one coefficient without building the table
  static double Coefficient(int j1, int j2, int m1, int m2, int j);

 private:
  int J1, J2;
  std::shared_ptr<const std::vector<double>> table;  // [(m1pos * (J2+1) + m2pos) * (J1+J2+1) + j]
};

// Lowest-Landau-level Coulomb pseudopotentials V_0..V_nbrFlux on the sphere in the units of
// EvaluatePseudopotentials (FQHE/src/Tools/FQHESpectrum/PseudoPotentials.cc:35) with zero
// layer separation, from the closed form
//   V_m = 2/sqrt(S) C(4S-2L, 2S-L) C(4S+2L+2, 2S+L+1) / C(4S+2, 2S+1)^2,  L = 2S - m.
std::vector<double> EvaluatePseudopotentials(int nbrFlux);
// FQHESpherePseudopotentialTools.cc:53: "Pseudopotentials = V_0 V_1 ..." in a key = value file
bool FQHESphereGetPseudopotentials(const std::string& fileName, int lzMax, std::vector<double>& pseudo);
// round like the 14 significant digits CoulombPseudopotentials writes
std::vector<double> RoundToFileDigits(const std::vector<double>& v);

// ------------------------------------------------------------------ architecture
// GPU counterpart of AbstractArchitecture (src/Architecture/AbstractArchitecture.h:45):
// owns the device context and routes the Hamiltonian-multiply operations to dh_hv.
class GPUArchitecture {
 public:
  explicit GPUArchitecture(int device = 0);
  ~GPUArchitecture();
  dh_context* Context() const { return ctx; }
  int GetArchitectureID() const { return 0x100; }  // new ArchitectureIDs entry "GPU"
  double* AllocateMemory(int64_t nbrDoubles);        // device vector storage
  void FreeMemory(double* p);
  void Check(int rc, const char* what) const;

 private:
  dh_context* ctx = nullptr;
};

// Device-resident block of k vectors in [row][k] layout (k = 1 is a RealVector).
class DeviceVector {
 public:
  DeviceVector() = default;
  DeviceVector(GPUArchitecture* arch, int64_t dim, int k = 1);
  ~DeviceVector();
  DeviceVector(const DeviceVector&) = delete;
  DeviceVector& operator=(const DeviceVector&) = delete;
  DeviceVector(DeviceVector&& o) noexcept;
  DeviceVector& operator=(DeviceVector&& o) noexcept;
  // a vector over device memory owned by the caller (released by the caller, not by this object)
  static DeviceVector View(GPUArchitecture* arch, double* d, int64_t dim, int k = 1);
  double* Data() const { return d; }
  int64_t GetVectorDimension() const { return dim; }
  int BlockSize() const { return k; }

 private:
  GPUArchitecture* arch = nullptr;
  double* d = nullptr;
  int64_t dim = 0;
  int k = 1;
  bool owns = true;
};

// ------------------------------------------------------------------ Hilbert spaces
class AbstractHilbertSpace {
 public:
  virtual ~AbstractHilbertSpace();
  int64_t GetHilbertSpaceDimension() const { return dh_basis_dimension(basis); }
  dh_basis* Basis() const { return basis; }

 protected:
  dh_basis* basis = nullptr;
};

// FermionOnSphere(nbrFermions, totalLz, lzMax) (FQHE/src/HilbertSpace/FermionOnSphere.cc:76)
class FermionOnSphere : public AbstractHilbertSpace {
 public:
  FermionOnSphere(GPUArchitecture* arch, int nbrFermions, int totalLz, int lzMax);
  int GetNbrParticles() const { return NbrFermions; }
  int GetLzMax() const { return LzMax; }
  int GetTotalLz() const { return TotalLz; }

 private:
  int NbrFermions, TotalLz, LzMax;
};

// Spin1_2Chain(chainLength, sz, memory) (Spin/src/HilbertSpace/Spin1_2Chain.cc:113)
class Spin1_2Chain : public AbstractHilbertSpace {
 public:
  Spin1_2Chain(GPUArchitecture* arch, int chainLength, int sz);
  int GetChainLength() const { return ChainLength; }

 private:
  int ChainLength, Sz;
};

// ------------------------------------------------------------------ Hamiltonians
// AbstractHamiltonian (src/Hamiltonian/AbstractHamiltonian.h:56) restricted to the
// operations of this path.  LowLevelAddMultiply adds H x to y for rows
// [firstComponent, firstComponent + nbrComponent).
class AbstractHamiltonian {
 public:
  virtual ~AbstractHamiltonian();
  int64_t GetHilbertSpaceDimension() const { return dh_hamiltonian_dimension(ham); }
  void ShiftHamiltonian(double shift);
  double GetShift() const { return Shift; }
  bool IsHermitian() const { return true; }
  DeviceVector& LowLevelAddMultiply(DeviceVector& vSource, DeviceVector& vDestination, int64_t firstComponent,
                                    int64_t nbrComponent);
  DeviceVector& LowLevelMultiply(DeviceVector& vSource, DeviceVector& vDestination, int64_t firstComponent,
                                 int64_t nbrComponent);
  // block products: the k vectors of a DeviceVector block share every index lookup
  DeviceVector& LowLevelMultipleAddMultiply(DeviceVector& vSources, DeviceVector& vDestinations);
  DeviceVector& LowLevelMultipleMultiply(DeviceVector& vSources, DeviceVector& vDestinations);
  // LowLevelMultipleAddMultiply(RealVector* vSources, RealVector* vDestinations, int nbrVectors,
  // int firstComponent, int nbrComponent) (src/Hamiltonian/AbstractHamiltonian.h:344): nbrVectors
  // separate one-vector DeviceVectors; groups of them share the index lookups (dh_hv_vectors)
  DeviceVector* LowLevelMultipleAddMultiply(DeviceVector* vSources, DeviceVector* vDestinations, int nbrVectors,
                                            int64_t firstComponent, int64_t nbrComponent);
  DeviceVector* LowLevelMultipleMultiply(DeviceVector* vSources, DeviceVector* vDestinations, int nbrVectors,
                                         int64_t firstComponent, int64_t nbrComponent);
  dh_hamiltonian* Handle() const { return ham; }
  GPUArchitecture* Architecture() const { return arch; }

 protected:
  GPUArchitecture* arch = nullptr;
  dh_hamiltonian* ham = nullptr;
  double Shift = 0.0;
};

// ParticleOnSphereGenericHamiltonian (FQHE/src/Hamiltonian/ParticleOnSphereGenericHamiltonian.cc:68)
// for fermions: two-body interaction from pseudopotentials V_0..V_lzMax.
class ParticleOnSphereGenericHamiltonian : public AbstractHamiltonian {
 public:
  ParticleOnSphereGenericHamiltonian(FermionOnSphere* particles, int nbrParticles, int lzmax,
                                     const std::vector<double>& pseudoPotential, GPUArchitecture* architecture);
  // the factor tables as uploaded (full list, pair-major, m3 ascending)
  const std::vector<double>& InteractionFactors() const { return Factors; }
  const std::vector<int>& PairOffsets() const { return PairOffset; }
  // adjoint channels: relative angular momenta m of the interaction split
  const std::vector<int>& ChannelMomenta() const { return ChannelM; }
  // g[c] = <psi|H_c|phi> per channel; if d_hpsi, also H psi
  void Bilinears(const double* d_psi, const double* d_phi, double* g, double* d_hpsi);

  // EvaluateInteractionFactors (:493) in the device layout; exposed for tests
  static void EvaluateInteractionFactors(int lzMax, const std::vector<double>& pseudo, std::vector<double>& factors,
                                         std::vector<int>& pairOffset);
  static void EvaluateChannelTable(int lzMax, std::vector<int>& channelM, std::vector<double>& cg);

 private:
  FermionOnSphere* Particles;
  int LzMax;
  std::vector<double> PseudoPotential, Factors, ChannelCG;
  std::vector<int> PairOffset, ChannelM;
};

// Total angular momentum of fermions on the sphere as a two-body operator in the factor layout
// of ParticleOnSphereGenericHamiltonian (the operator of ParticleOnSphereL2Hamiltonian,
// FQHE/src/Hamiltonian/ParticleOnSphereL2Hamiltonian.cc:240): l2Factor * L^2 on the sector of
// nbrParticles fermions with twice the total Lz equal to totalLz.  With pseudopotentials it is
// added to their interaction (the --l2-factor option of the reference programs).
class ParticleOnSphereL2Hamiltonian : public AbstractHamiltonian {
 public:
  ParticleOnSphereL2Hamiltonian(FermionOnSphere* particles, int nbrParticles, int lzmax, double l2Factor,
                                GPUArchitecture* architecture, const std::vector<double>* pseudoPotential = nullptr);
  static void EvaluateInteractionFactors(int lzMax, int nbrParticles, int totalLz, double l2Factor,
                                         std::vector<double>& factors, std::vector<int>& pairOffset, double& shift);

 private:
  FermionOnSphere* Particles;
  std::vector<double> Factors;
  std::vector<int> PairOffset;
};

// Three-body term of AbstractQHEOnSphereNBodyInteractionHamiltonian
// (FQHE/src/Hamiltonian/AbstractQHEOnSphereNBodyInteractionHamiltonian.cc:107) with explicit MN
// factors: H = sum_S sum_{i,l} F_S[i][l] ProdAd(M_l) ProdA(N_i), the tuples of each orbital sum S
// given in descending orbital order and F_S symmetric.
class ParticleOnSphereNBodyHamiltonian : public AbstractHamiltonian {
 public:
  ParticleOnSphereNBodyHamiltonian(FermionOnSphere* particles, int nbody, const std::vector<int>& sumCount,
                                   const std::vector<int>& tuples, const std::vector<double>& factors,
                                   GPUArchitecture* architecture);

 private:
  FermionOnSphere* Particles;
  int NBody;
};

// SpinChainHamiltonian (Spin/src/Hamiltonian/SpinChainHamiltonian.cc:97)
class SpinChainHamiltonian : public AbstractHamiltonian {
 public:
  SpinChainHamiltonian(Spin1_2Chain* chain, int nbrSpin, const std::vector<double>& j, const std::vector<double>& jz,
                       bool periodicBoundaryConditions, GPUArchitecture* architecture);
};

// ------------------------------------------------------------------ operations
// VectorHamiltonianMultiplyOperation (src/Architecture/ArchitectureOperation/VectorHamiltonianMultiplyOperation.cc):
// on the GPU architecture the whole row range runs as one device launch.
class VectorHamiltonianMultiplyOperation {
 public:
  VectorHamiltonianMultiplyOperation(AbstractHamiltonian* hamiltonian, DeviceVector* source, DeviceVector* destination);
  bool ApplyOperation(GPUArchitecture* architecture);

 private:
  AbstractHamiltonian* Hamiltonian;
  DeviceVector *Source, *Destination;
};

// MultipleVectorHamiltonianMultiplyOperation
// (src/Architecture/ArchitectureOperation/MultipleVectorHamiltonianMultiplyOperation.cc): H applied
// to nbrVectors separate vectors; on the GPU architecture the whole row range is one call of
// LowLevelMultipleMultiply (RawApplyOperation :427 with the full range).
class MultipleVectorHamiltonianMultiplyOperation {
 public:
  MultipleVectorHamiltonianMultiplyOperation(AbstractHamiltonian* hamiltonian, DeviceVector* sources,
                                             DeviceVector* destinations, int nbrVectors);
  void SetIndicesRange(int64_t firstComponent, int64_t nbrComponent);
  bool ApplyOperation(GPUArchitecture* architecture);

 private:
  AbstractHamiltonian* Hamiltonian;
  DeviceVector *Sources, *Destinations;
  int NbrVectors;
  int64_t FirstComponent = 0, NbrComponent = -1;
};

// ------------------------------------------------------------------ Lanczos
struct LanczosResult {
  std::vector<double> eigenvalues;
  int iterations = 0;        // vector products (block: blocks * k)
  int blockIterations = 0;   // Lanczos steps
  bool converged = false;
  double seconds = 0.0;
};

// FullReorthogonalizedLanczosAlgorithm (src/LanczosAlgorithm/FullReorthogonalizedLanczosAlgorithm.cc)
// and FullReorthogonalizedBlockLanczosAlgorithm (FullReorthogonalizedBlockLanczosAlgorithm.cc):
// block size 1 is the plain algorithm.  The Krylov basis stays on the device; every new block is
// orthogonalised against all previous ones (two classical Gram-Schmidt passes).  Convergence as
// TestConvergence (:325, strong form): every wanted eigenvalue changes by less than
// precision * |value| between two consecutive checks.
class FullReorthogonalizedBlockLanczosAlgorithm {
 public:
  FullReorthogonalizedBlockLanczosAlgorithm(GPUArchitecture* architecture, AbstractHamiltonian* hamiltonian,
                                            int nbrEigenvalue, int blockSize, int maxIter, double precision = 1e-14,
                                            double residual = 0.0);
  ~FullReorthogonalizedBlockLanczosAlgorithm();
  void InitializeLanczosAlgorithm(uint64_t seed);
  // FullReorthogonalizedLanczosAlgorithmWithDiskStorage: from now on every Krylov vector is also
  // written as directory/vector.<i> (DiagHam vector files; column p of block j is vector j*k+p)
  // and the recurrence state as directory/lanczos.dat, after each step.
  void SetDiskStorage(const std::string& directory);
  // ResumeLanczosAlgorithm (FullReorthogonalizedLanczosAlgorithmWithDiskStorage.cc:176): reload
  // the state and the Krylov vectors that SetDiskStorage wrote and continue from there; the
  // resumed run repeats the arithmetic of an uninterrupted one.
  void ResumeLanczosAlgorithm();
  void RunLanczosAlgorithm(int nbrIter);
  bool TestConvergence();
  std::vector<double> GetEigenvalues() const { return Current; }
  // Ritz vectors of the nbr lowest eigenvalues into d_out (nbr * dim doubles, vector-major)
  void GetEigenstates(int nbr, double* d_out);
  LanczosResult Solve();
  int NbrBlocks() const;

 private:
  struct HamiltonianOps;
  void Diagonalize(std::vector<double>& evals, std::vector<double>* evecs);
  bool Dense() const;
  void DenseSolve();
  void WriteState();
  void WriteBlock(int j);
  std::string DiskDirectory;
  GPUArchitecture* Arch;
  AbstractHamiltonian* Hamiltonian;
  int NbrEigenvalue, BlockSize, MaxIter;
  double Precision, Residual;
  int64_t Dim;
  std::vector<double*> d_basis;  // chunks of Krylov storage (the first ends with the scratch block)
  int MaxBlocks;
  std::unique_ptr<BlockOps> Ops;
  std::unique_ptr<SegmentedBlockLanczos<double>> Engine;
  std::vector<double> Current;
};

// All Lz sectors of one fermions-on-the-sphere problem solved together: the per-sector loop of
// FQHESphereFermionsTwoBodyGeneric.cc:270 + QHEOnSphereMainTask becomes one block Lanczos per
// sector whose steps advance every unconverged sector at once (one persistent H.v launch and
// segmented reorthogonalisation per step).  Sectors too small for a Krylov block are
// diagonalised densely.  Sectors are processed in waves whose Krylov bases fit `memoryBytes`.
struct SectorResult {
  int totalLz = 0;
  int64_t dimension = 0;
  std::vector<double> eigenvalues;
  int products = 0;   // vector products over every attempt
  bool converged = false;
  int blocks = 0;     // Krylov blocks of the last attempt
  bool deferred = false;  // the last attempt stopped because the device could not hold its storage
};

// A Krylov storage growth that the device memory cannot hold (GrowKrylovStorage).
struct KrylovStorageFull : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Progress lines of long solves on stderr when the environment sets DHH_VERBOSE (like the
// --show-itertime option of the reference programs): at most one per `seconds` per caller.
class ProgressLog {
 public:
  explicit ProgressLog(double seconds = 20.0)
      : On(std::getenv("DHH_VERBOSE") != nullptr), Every(seconds), Last(std::chrono::steady_clock::now()) {}
  bool Due() {
    if (!On) return false;
    auto now = std::chrono::steady_clock::now();
    if (std::chrono::duration<double>(now - Last).count() < Every) return false;
    Last = now;
    return true;
  }
  bool Enabled() const { return On; }

 private:
  bool On;
  double Every;
  std::chrono::steady_clock::time_point Last;
};

// Launch shape of the products of a batched sector solve (dh_sector_set_launch and
// dh_sector_set_row_mode); lanesPerRow == 0 keeps the solver's built-in choice.
struct SectorLaunch {
  int lanesPerRow = 0, blockSize = 256, rowMode = DH_INDEX_RANK;
};

// Rows from which a sector of a batched sphere solve runs alone, through the single-space solver
// (0: every sector in the waves).  Such a sector fills the device by itself, so batching saves no
// per-step overhead on it, and in a wave its vectors compete for the caches with its neighbours':
// the all-sector product of N=12 2S=33 makes 4.7e11 elements/s against 5.6e11 for the Lz = 0
// sector alone, and the 133 sectors took 235 s batched against 198 s one after the other.
constexpr int64_t kSectorAloneRows = 2000000;

// free device memory of an architecture (0 if it cannot be read)
inline size_t DeviceFreeBytes(GPUArchitecture* arch) {
  size_t freeBytes = 0, totalBytes = 0;
  return dh_memory_info(arch->Context(), &freeBytes, &totalBytes) == DH_OK ? freeBytes : 0;
}

// BlockOps over the rows of a sector set: k vectors per block, kr doubles per row (k for real
// vectors, 2k for complex ones)
class SectorSetOps : public BlockOps {
 public:
  SectorSetOps(GPUArchitecture* arch, dh_sector_set* set, int k, int kr, int64_t rows);
  void Product(const double* X, double* W, const unsigned char* active) override;
  void Dots(const double* V, int nb, const double* W, const unsigned char* active, double* out) override;
  void Axpy(const double* V, int nb, double* W, const unsigned char* active, const double* C) override;
  void Rmul(double* W, const unsigned char* active, const double* M) override;
  void Random(double* W, uint64_t seed) override;

 private:
  GPUArchitecture* Arch;
  dh_sector_set* Set;
  int K, KR;
  int64_t Rows;
};

// A sector set and its Krylov storage, released on every way out of a solve (errors leave as
// exceptions, and the storage pointer changes when it grows).
struct SectorSolveResources {
  explicit SectorSolveResources(GPUArchitecture* a) : arch(a) {}
  ~SectorSolveResources() {
    dh_synchronize(arch->Context());
    for (double* p : basis) arch->FreeMemory(p);
    if (set) dh_sector_set_destroy(set);
  }
  SectorSolveResources(const SectorSolveResources&) = delete;
  SectorSolveResources& operator=(const SectorSolveResources&) = delete;
  GPUArchitecture* arch;
  std::vector<double*> basis;  // the chunks of the Krylov storage
  dh_sector_set* set = nullptr;
};

// Krylov blocks a basis starts with (about 256 vectors, never more than maxBlocks); the engine
// grows it by chunks when it is full.  Holding all maxBlocks + 2 blocks from the start (160 GB for
// 3000 vectors over 6.7 M rows) made every later solve of the same process run its
// reorthogonalisation at about half speed on MI355X; sized to what the run uses, it does not.
inline int InitialKrylovBlocks(int k, int maxBlocks) { return std::min(maxBlocks, std::max(32, 256 / k)); }
// a further chunk of Krylov storage, `blocks` blocks of blockDoubles: `blocks` is lowered (not
// below `need`) to what the device's free memory holds, keeping 1 GiB spare; when not even `need`
// blocks fit, KrylovStorageFull is thrown.  The chunk is appended to `chunks`, whose owner frees them.
double* GrowKrylovStorage(GPUArchitecture* arch, std::vector<double*>& chunks, int& blocks, int need,
                          int64_t blockDoubles);

// Krylov blocks each sector may use (maxIter / k + 1, and no more than its space needs); returns
// the largest
inline int SectorBlockCaps(const std::vector<int64_t>& dims, int k, int maxIter, std::vector<int>& caps) {
  int most = 2;
  caps.resize(dims.size());
  for (size_t i = 0; i < dims.size(); ++i) {
    caps[i] = (int)std::min<int64_t>(std::max(3, maxIter / k + 1), (dims[i] + k - 1) / k + 1);
    most = std::max(most, caps[i]);
  }
  return most;
}

// Steps every unconverged sector of `lanczos` together until each has converged (strong test of
// FullReorthogonalizedLanczosAlgorithm.cc:325), completed its Krylov space or used its blocks.
// The product of the next step is enqueued before the convergence tests of this one run on the
// host, for every sector that can still step; a sector that converges in those tests drops out of
// the orthogonalisation, and its one extra product is discarded.
//
// When the Krylov storage cannot grow any more (KrylovStorageFull), the sectors still running after
// this step's convergence tests stop with `deferred` set, for the caller to solve again with more
// room; the others keep their results.
template <class T>
void RunSectorLanczos(SegmentedBlockLanczos<T>& lanczos, const std::vector<int>& blockCap, int nbrEigenvalue,
                      double precision, uint64_t seed, std::vector<SectorResult*>& results) {
  const int na = (int)results.size();
  const int k = lanczos.BlockSize();
  for (SectorResult* r : results) {
    r->converged = false;
    r->deferred = false;
    r->blocks = 0;
  }
  lanczos.Start(seed);
  std::vector<unsigned char> active(na, 1), next(na, 0);
  std::vector<char> done(na, 0);
  lanczos.LaunchProduct(active);
  bool full = false;
  ProgressLog log;
  while (!full) {
    std::vector<int> act;
    for (int i = 0; i < na; ++i)
      if (active[i]) act.push_back(i);
    if (act.empty()) break;
    if (log.Due())
      std::fprintf(stderr, "[sector Lanczos] step %d, %zu of %d sectors running\n", lanczos.Steps(), act.size(), na);
    lanczos.FinishStep(active);
    bool any = false;
    for (int i = 0; i < na; ++i) {
      next[i] = active[i] && !(lanczos.Exhausted(i) || lanczos.NbrBlocks(i) >= blockCap[i]);
      any = any || next[i];
    }
    if (any) {
      try {
        lanczos.LaunchProduct(next);
      } catch (const KrylovStorageFull&) {
        full = true;
      }
    }
    for (int i : act) {
      results[i]->products += k;
      results[i]->blocks = lanczos.NbrBlocks(i);
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t a = 0; a < act.size(); ++a) {
      const int i = act[a];
      const bool last = !next[i];
      if (!last && !full && !lanczos.CheckDue(i, nbrEigenvalue)) continue;
      std::vector<double> cur;
      const bool ok = lanczos.TestConvergence(i, nbrEigenvalue, precision, cur);
      results[i]->eigenvalues = cur;
      if (ok) results[i]->converged = true;
      if (ok || last) done[i] = 1;
    }
    for (int i = 0; i < na; ++i) active[i] = next[i] && !done[i];
  }
  if (full)
    for (int i = 0; i < na; ++i)
      if (active[i]) results[i]->deferred = true;
}

// Krylov blocks a sector of `dim` rows is expected to need for nbrEigenvalue eigenvalues at block
// size k, within the sector's own space and maxIter.  Measured on MI355X runs of Coulomb sectors
// for 10 eigenvalues at k = 1: 229 vectors on average over N=10 2S=27, up to 475 for the largest
// sectors of N=12 2S=33; the estimate covers the large ones, 40 vectors per wanted eigenvalue plus
// 100, each further vector of a block adding 30 %.  An underestimate is costly: a sector whose
// storage cannot grow in a full wave stops and runs again from the start in a later wave.
inline int ExpectedKrylovBlocks(int64_t dim, int k, int nbrEigenvalue, int maxIter) {
  const double vectors = (40.0 * nbrEigenvalue + 100.0) * (1.0 + 0.3 * (k - 1));
  const int64_t blocks = (int64_t)std::ceil(vectors / k);
  return (int)std::max<int64_t>(3, std::min<int64_t>({blocks, (dim + k - 1) / k + 1, maxIter / k + 1}));
}

// Waves of a batched sector solve.  Sectors are taken largest first and admitted to a wave while its
// Krylov storage with room to grow, (1.5 x expected blocks + 2) x rows x kr doubles plus the basis
// words, fits the smaller of memoryBytes and the device's free memory less 2 GiB.  A wave's storage starts small
// and grows as the run needs it (SegmentedBlockLanczos::SetGrowth); sectors stopped because it
// could not grow go back to the queue expecting 1.5 times the blocks they reached, and after each
// wave the expectation follows the blocks its sectors used.  solveWave(sectors, blocks) fills out[s]
// of the sectors it is given (blocks: the expected Krylov blocks of the wave, the storage to start
// with); freeBytes() is the device memory free before a wave.  Returns the number of waves.
template <class F, class M>
int RunSectorWaves(const std::vector<int64_t>& dims, int k, int kr, int nbrEigenvalue, int maxIter, size_t memoryBytes,
                   std::vector<SectorResult>& out, F solveWave, M freeBytes) {
  std::vector<int> queue;
  for (size_t s = 0; s < dims.size(); ++s) queue.push_back((int)s);
  std::stable_sort(queue.begin(), queue.end(), [&](int a, int b) { return dims[a] > dims[b]; });
  std::vector<int> expect(dims.size());
  for (size_t s = 0; s < dims.size(); ++s) expect[s] = ExpectedKrylovBlocks(dims[s], k, nbrEigenvalue, maxIter);
  int waves = 0;
  while (!queue.empty()) {
    const size_t avail = freeBytes(), spare = (size_t)2 << 30;
    const size_t budget = std::min(memoryBytes, avail > spare ? avail - spare : (size_t)0);
    std::vector<int> wave, rest;
    int blocks = 0, room = 0;
    int64_t rows = 0;
    for (int s : queue) {
      const int cap = (int)std::min<int64_t>((dims[s] + k - 1) / k + 1, maxIter / k + 1);
      const int b = std::max(blocks, std::min(expect[s], cap));
      // admitted with room for half as many blocks again: the storage of a wave is shared by its
      // sectors (block-major), so when one sector outgrows a full device every sector still running
      // in that wave stops and starts over later; with the room, growth past the expectation is
      // the rule and restarts the exception (N=10 2S=27 sectors need 0.5-1.1 x the 500 blocks
      // expected for 10 eigenvalues at k = 1)
      const int r = std::max(room, std::min(b + (b + 1) / 2, std::max(b, cap)));
      const size_t need = sizeof(double) * ((size_t)(r + 2) * (size_t)(rows + dims[s]) * kr + (size_t)(rows + dims[s]));
      if (!wave.empty() && need > budget) {
        rest.push_back(s);
        continue;
      }
      wave.push_back(s);
      blocks = b;
      room = r;
      rows += dims[s];
    }
    if (std::getenv("DHH_VERBOSE"))
      std::fprintf(stderr, "[sector waves] wave %d: %zu sectors, %lld rows, %d expected blocks, %zu of %zu queued\n",
                   waves + 1, wave.size(), (long long)rows, blocks, wave.size(), queue.size());
    solveWave(wave, blocks);
    ++waves;
    int reached = 0;
    std::vector<int> again;
    for (int s : wave) {
      reached = std::max(reached, out[s].blocks);
      if (out[s].deferred) again.push_back(s);
    }
    if (!again.empty() && again.size() == wave.size() && wave.size() == 1)
      throw std::runtime_error("a sector needs more Krylov storage than the device holds");
    for (int s : again) expect[s] = std::max(expect[s] + 1, (out[s].blocks * 3 + 1) / 2);
    // the sectors left are no larger than this wave's: expect at most what it used, with margin
    if (again.size() < wave.size())
      for (int s : rest) expect[s] = std::min(expect[s], std::max(3, (reached * 23 + 19) / 20));
    queue = again;
    queue.insert(queue.end(), rest.begin(), rest.end());
    std::stable_sort(queue.begin(), queue.end(), [&](int a, int b) { return dims[a] > dims[b]; });
  }
  return waves;
}

class MultiSectorBlockLanczos {
 public:
  MultiSectorBlockLanczos(GPUArchitecture* architecture, int nbrFermions, int lzMax, const std::vector<int>& totalLz,
                          const std::vector<double>& pseudoPotential, int nbrEigenvalue, int blockSize, int maxIter,
                          double precision = 1e-14, size_t memoryBytes = 0);
  std::vector<SectorResult> Solve(uint64_t seed, double* seconds = nullptr);
  int NbrWaves() const { return Waves; }
  void SetLaunch(const SectorLaunch& launch) { Launch = launch; }
  // rows from which a sector runs alone (kSectorAloneRows by default; 0: none)
  void SetAloneRows(int64_t rows) { AloneRows = rows; }

 private:
  // expectedBlocks: the Krylov blocks the wave planner expects the sectors to need
  void SolveWave(const std::vector<int>& sectors, uint64_t seed, std::vector<SectorResult>& out, int expectedBlocks);
  SectorLaunch Launch;
  int64_t AloneRows = kSectorAloneRows;
  GPUArchitecture* Arch;
  int NbrFermions, LzMax, NbrEigenvalue, BlockSize, MaxIter;
  double Precision;
  size_t MemoryBytes;
  int Waves = 0;
  std::vector<int> TotalLz;
  std::vector<double> Pseudo, Factors;
  std::vector<int> PairOffset;
};

// Tridiagonal QL with implicit shifts (RealTriDiagonalSymmetricMatrix.cc:687 semantics):
// eigenvalues ascending; optional eigenvectors (column-major n x n) of the tridiagonal.
// Eigenvector (normalised) of the symmetric tridiagonal (d, e) for its eigenvalue theta, by inverse
// iteration; O(n).
std::vector<double> TridiagonalEigenvector(const std::vector<double>& d, const std::vector<double>& e, double theta);
void TridiagonalEigen(int n, std::vector<double> d, std::vector<double> e, std::vector<double>& evals,
                      std::vector<double>* evecs);
// ------------------------------------------------------------------ Chern band (momentum space)
// Lower band of TightBindingModelCheckerboardLattice as OneBodyBasis[k][0][alpha] (u[2k + alpha])
// and the band energies (energies[2k + band]).
void CheckerboardLowerBand(int nx, int ny, double t1, double t2, double tpp, double mus, double gx, double gy,
                           std::vector<std::complex<double>>& u, std::vector<double>& energies);
// ParticleOnLatticeCheckerboardLatticeSingleBandHamiltonian factors in the layout of
// dh_sector_set_lattice_factors (complex entries interleaved in `factors`).
void CheckerboardSingleBandFactors(int nx, int ny, const std::vector<std::complex<double>>& u, double uPotential,
                                   double vPotential, bool flatBand, std::vector<int>& sumPairCount,
                                   std::vector<int>& sumPairs, std::vector<double>& factors);

// Two-body factors of fermions on a torus of nbrFlux flux quanta and aspect ratio `ratio`
// (ParticleOnTorusGenericHamiltonian with pseudopotentials when coulombLevel < 0,
// ParticleOnTorusCoulombHamiltonian in Landau level coulombLevel otherwise), in the layout of
// dh_sector_set_lattice_factors for a momentum sector set with Nx = 1, Ny = nbrFlux (orbital = ky).
void TorusInteractionFactors(int nbrFlux, double ratio, const std::vector<double>& pseudo, int coulombLevel,
                             std::vector<int>& sumPairCount, std::vector<int>& sumPairs,
                             std::vector<double>& factors);

// FCICheckerboardLatticeModel's (kx, ky) loop (FTI/src/Programs/FCI/FCICheckerboardLatticeModel.cc:412)
// with FullReorthogonalizedComplex(Block)LanczosAlgorithm per sector, all sectors advanced together.
class LatticeSectorsBlockLanczos {
 public:
  LatticeSectorsBlockLanczos(GPUArchitecture* architecture, int nbrFermions, int nx, int ny,
                             const std::vector<int>& kx, const std::vector<int>& ky,
                             const std::vector<int>& sumPairCount, const std::vector<int>& sumPairs,
                             const std::vector<double>& factors, int nbrEigenvalue, int blockSize, int maxIter,
                             double precision = 1e-14, size_t memoryBytes = 0);
  // the sectors are (Kx, Ky) sectors of a torus with magnetic translations
  // (FermionOnTorusWithMagneticTranslations; nx = 1, ny = Nphi, factors of the torus)
  void UseTorusTranslations() { TorusTranslations = true; }
  std::vector<SectorResult> Solve(uint64_t seed, double* seconds = nullptr);
  int NbrWaves() const { return Waves; }
  void SetLaunch(const SectorLaunch& launch) { Launch = launch; }

 private:
  // expectedBlocks: the Krylov blocks the wave planner expects the sectors to need
  void SolveWave(const std::vector<int>& sectors, uint64_t seed, std::vector<SectorResult>& out, int expectedBlocks);
  dh_sector_set* CreateSet(const std::vector<int>& kx, const std::vector<int>& ky) const;
  bool TorusTranslations = false;
  SectorLaunch Launch;
  GPUArchitecture* Arch;
  int NbrFermions, NbrSiteX, NbrSiteY, NbrEigenvalue, BlockSize, MaxIter;
  double Precision;
  size_t MemoryBytes;
  int Waves = 0;
  std::vector<int> Kx, Ky, SumPairCount, SumPairs;
  std::vector<double> Factors;
};

// DiagHam binary vector files (RealVector::WriteVector src/Vector/RealVector.cc:3669,
// ComplexVector::WriteVector src/Vector/ComplexVector.cc:2606, RealVector::ReadVector :3798) from
// and into device memory; complex vectors are interleaved (re, im).
void WriteDeviceVector(GPUArchitecture* arch, const char* fileName, const double* d_vec, int64_t dim, bool complex);
void ReadDeviceVector(GPUArchitecture* arch, const char* fileName, double* d_vec, int64_t dim, bool complex);
// The dimension in the file's header; headerBytes (if given) receives the header's length (4 for
// the int32 form, 12 for -1 and an int64).
int64_t VectorFileDimension(const char* fileName, bool complex, int64_t* headerBytes = nullptr);

// Symmetric band matrix (half-bandwidth b, dense row-major n x n) -> tridiagonal by Givens
// rotations with bulge chasing; returns diagonal and off-diagonal, and accumulates Q if given.
void BandToTridiagonal(int n, int b, std::vector<double>& a, std::vector<double>& d, std::vector<double>& e,
                       std::vector<double>* q);

}  // namespace dhh
