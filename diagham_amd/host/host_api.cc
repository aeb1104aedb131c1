// C facade of the host layer (include/diagham_host.h).
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "../../include/diagham_host.h"
#include "diagham_host.hpp"

using namespace dhh;

// An architecture is reference counted: dhh_arch_create hands its caller one reference and every
// problem holds one more, so the device context outlives the problems built on it whatever order
// their owners release them in.  dhh_arch_destroy gives up the caller's reference.
struct dhh_arch {
  std::unique_ptr<GPUArchitecture> arch;
  std::atomic<int> refs{1};
  bool closed = false;
};

static void arch_release(dhh_arch* a) {
  if (a->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) delete a;
}

// counted reference of a problem to its architecture (released after the problem's space and
// Hamiltonian, which are declared after it)
class ArchRef {
 public:
  ArchRef() = default;
  ArchRef(const ArchRef&) = delete;
  ArchRef& operator=(const ArchRef&) = delete;
  ~ArchRef() {
    if (p) arch_release(p);
  }
  ArchRef& operator=(dhh_arch* a) {
    if (a) a->refs.fetch_add(1, std::memory_order_relaxed);
    if (p) arch_release(p);
    p = a;
    return *this;
  }
  dhh_arch* operator->() const { return p; }

 private:
  dhh_arch* p = nullptr;
};

struct dhh_problem {
  ArchRef arch;
  std::unique_ptr<AbstractHilbertSpace> space;
  std::unique_ptr<AbstractHamiltonian> hamiltonian;
  std::vector<int> channels;
};

static thread_local std::string g_error;
static thread_local int g_lastWaves = 0;
static thread_local SectorLaunch g_sectorLaunch;  // dhh_sectors_launch
static thread_local int64_t g_aloneRows = kSectorAloneRows;  // dhh_sectors_alone_rows

template <typename F>
static int guarded(F&& f) {
  try {
    return f();
  } catch (const std::invalid_argument& e) {
    g_error = e.what();
    return DH_E_ARG;
  } catch (const std::exception& e) {
    g_error = e.what();
    return DH_E_STATE;
  }
}

extern "C" const char* dhh_last_error(void) { return g_error.c_str(); }

extern "C" int dhh_last_solve_waves(void) { return g_lastWaves; }

extern "C" int dhh_sectors_launch(int lanesPerRow, int blockSize, int rowMode) {
  if (lanesPerRow == 0) {
    g_sectorLaunch = SectorLaunch();
    return DH_OK;
  }
  if ((lanesPerRow != 1 && lanesPerRow != 2 && lanesPerRow != 4 && lanesPerRow != 8) || blockSize < 64 ||
      blockSize > 1024 || blockSize % 64 ||
      (rowMode != DH_INDEX_RANK && rowMode != DH_INDEX_RANK_FREE && rowMode != DH_INDEX_RANK_RANGE)) {
    g_error = "launch shape: lanes per row in {1, 2, 4, 8}, workgroup a multiple of 64 up to 1024, "
              "row mode DH_INDEX_RANK, DH_INDEX_RANK_FREE or DH_INDEX_RANK_RANGE";
    return DH_E_ARG;
  }
  g_sectorLaunch.lanesPerRow = lanesPerRow;
  g_sectorLaunch.blockSize = blockSize;
  g_sectorLaunch.rowMode = rowMode;
  return DH_OK;
}

extern "C" int dhh_sectors_alone_rows(int64_t rows) {
  if (rows < 0) {
    g_error = "alone rows: 0 (every sector batched) or a positive row count";
    return DH_E_ARG;
  }
  g_aloneRows = rows;
  return DH_OK;
}

extern "C" int dhh_arch_create(dhh_arch** out, int device) {
  if (!out) return DH_E_ARG;
  return guarded([&] {
    auto* a = new dhh_arch();
    a->arch.reset(new GPUArchitecture(device));
    *out = a;
    return DH_OK;
  });
}

extern "C" int dhh_arch_destroy(dhh_arch* a) {
  if (!a) return DH_E_ARG;
  if (a->closed) {
    g_error = "architecture already destroyed";
    return DH_E_STATE;
  }
  a->closed = true;
  arch_release(a);
  return DH_OK;
}

extern "C" dh_context* dhh_arch_context(dhh_arch* a) { return a ? a->arch->Context() : nullptr; }

extern "C" int dhh_sphere_problem_create(dhh_arch* a, dhh_problem** out, int n, int totalLz, int lzMax,
                                         const double* pseudo) {
  if (!a || !out || !pseudo) return DH_E_ARG;
  if (a->closed) {
    g_error = "architecture destroyed";
    return DH_E_STATE;
  }
  return guarded([&] {
    auto p = std::make_unique<dhh_problem>();
    p->arch = a;
    auto* space = new FermionOnSphere(a->arch.get(), n, totalLz, lzMax);
    p->space.reset(space);
    std::vector<double> v(pseudo, pseudo + lzMax + 1);
    auto* h = new ParticleOnSphereGenericHamiltonian(space, n, lzMax, v, a->arch.get());
    p->hamiltonian.reset(h);
    p->channels = h->ChannelMomenta();
    *out = p.release();
    return DH_OK;
  });
}

extern "C" int dhh_write_vector(dhh_arch* a, const char* fileName, const double* d_vec, int64_t dim, int complex) {
  if (!a || !fileName || (!d_vec && dim > 0)) return DH_E_ARG;
  return guarded([&] {
    WriteDeviceVector(a->arch.get(), fileName, d_vec, dim, complex != 0);
    return DH_OK;
  });
}

extern "C" int dhh_read_vector(dhh_arch* a, const char* fileName, double* d_vec, int64_t dim, int complex) {
  if (!a || !fileName || (!d_vec && dim > 0)) return DH_E_ARG;
  return guarded([&] {
    ReadDeviceVector(a->arch.get(), fileName, d_vec, dim, complex != 0);
    return DH_OK;
  });
}

extern "C" int dhh_vector_file_dimension(const char* fileName, int complex, int64_t* dim) {
  if (!fileName || !dim) return DH_E_ARG;
  return guarded([&] {
    *dim = VectorFileDimension(fileName, complex != 0);
    return DH_OK;
  });
}

extern "C" int dhh_sphere_l2_problem_create(dhh_arch* a, dhh_problem** out, int n, int totalLz, int lzMax,
                                            const double* pseudo, double l2Factor) {
  if (!a || !out) return DH_E_ARG;
  if (a->closed) {
    g_error = "architecture destroyed";
    return DH_E_STATE;
  }
  return guarded([&] {
    auto p = std::make_unique<dhh_problem>();
    p->arch = a;
    auto* space = new FermionOnSphere(a->arch.get(), n, totalLz, lzMax);
    p->space.reset(space);
    std::vector<double> v;
    if (pseudo) v.assign(pseudo, pseudo + lzMax + 1);
    p->hamiltonian.reset(
        new ParticleOnSphereL2Hamiltonian(space, n, lzMax, l2Factor, a->arch.get(), pseudo ? &v : nullptr));
    *out = p.release();
    return DH_OK;
  });
}

extern "C" int dhh_sphere_l2_factors(int lzMax, int n, int totalLz, double l2Factor, double* factors,
                                     int* pairOffset, int* nbr, double* shift) {
  if (!factors || !pairOffset || !nbr || !shift) return DH_E_ARG;
  return guarded([&] {
    std::vector<double> f;
    std::vector<int> off;
    ParticleOnSphereL2Hamiltonian::EvaluateInteractionFactors(lzMax, n, totalLz, l2Factor, f, off, *shift);
    if ((int)f.size() > *nbr) return DH_E_RANGE;
    std::memcpy(factors, f.data(), sizeof(double) * f.size());
    std::memcpy(pairOffset, off.data(), sizeof(int) * off.size());
    *nbr = (int)f.size();
    return DH_OK;
  });
}

extern "C" int dhh_sphere_nbody_problem_create(dhh_arch* a, dhh_problem** out, int n, int totalLz, int lzMax,
                                               int nbody, const int* sumCount, int nbrSums, const int* tuples,
                                               const double* factors) {
  if (!a || !out || !sumCount || !tuples || !factors || nbrSums <= 0 || nbody <= 0) return DH_E_ARG;
  if (a->closed) {
    g_error = "architecture destroyed";
    return DH_E_STATE;
  }
  return guarded([&] {
    auto p = std::make_unique<dhh_problem>();
    p->arch = a;
    auto* space = new FermionOnSphere(a->arch.get(), n, totalLz, lzMax);
    p->space.reset(space);
    std::vector<int> counts(sumCount, sumCount + nbrSums);
    int64_t nt = 0, nf = 0;
    for (int c : counts) {
      nt += c;
      nf += (int64_t)c * c;
    }
    std::vector<int> t(tuples, tuples + nt * nbody);
    std::vector<double> f(factors, factors + nf);
    p->hamiltonian.reset(new ParticleOnSphereNBodyHamiltonian(space, nbody, counts, t, f, a->arch.get()));
    *out = p.release();
    return DH_OK;
  });
}

extern "C" int dhh_spin_problem_create(dhh_arch* a, dhh_problem** out, int L, int sz, const double* j,
                                       const double* jz, int periodic) {
  if (!a || !out || !j || !jz) return DH_E_ARG;
  if (a->closed) {
    g_error = "architecture destroyed";
    return DH_E_STATE;
  }
  return guarded([&] {
    auto p = std::make_unique<dhh_problem>();
    p->arch = a;
    auto* chain = new Spin1_2Chain(a->arch.get(), L, sz);
    p->space.reset(chain);
    std::vector<double> J(j, j + L), Jz(jz, jz + L);
    p->hamiltonian.reset(new SpinChainHamiltonian(chain, L, J, Jz, periodic != 0, a->arch.get()));
    *out = p.release();
    return DH_OK;
  });
}

extern "C" int dhh_problem_destroy(dhh_problem* p) {
  if (!p) return DH_E_ARG;
  p->hamiltonian.reset();
  p->space.reset();
  delete p;
  return DH_OK;
}

extern "C" dh_basis* dhh_problem_basis(dhh_problem* p) { return p ? p->space->Basis() : nullptr; }
extern "C" dh_hamiltonian* dhh_problem_hamiltonian(dhh_problem* p) { return p ? p->hamiltonian->Handle() : nullptr; }

extern "C" int dhh_problem_multiple_multiply(dhh_problem* p, const double* const* x, double* const* y, int n,
                                             int add, int64_t first, int64_t count) {
  if (!p || !x || !y || n < 0) return DH_E_ARG;
  return guarded([&] {
    GPUArchitecture* arch = p->arch->arch.get();
    const int64_t dim = p->hamiltonian->GetHilbertSpaceDimension();
    std::vector<DeviceVector> src, dst;
    src.reserve(n);
    dst.reserve(n);
    for (int v = 0; v < n; ++v) {
      src.push_back(DeviceVector::View(arch, const_cast<double*>(x[v]), dim));
      dst.push_back(DeviceVector::View(arch, y[v], dim));
    }
    if (add)
      p->hamiltonian->LowLevelMultipleAddMultiply(src.data(), dst.data(), n, first, count);
    else {
      MultipleVectorHamiltonianMultiplyOperation op(p->hamiltonian.get(), src.data(), dst.data(), n);
      op.SetIndicesRange(first, count);
      op.ApplyOperation(arch);
    }
    return DH_OK;
  });
}

extern "C" int dhh_problem_channels(dhh_problem* p, int* momenta, int capacity) {
  if (!p) return DH_E_ARG;
  int n = (int)p->channels.size();
  for (int i = 0; i < n && i < capacity; ++i) momenta[i] = p->channels[i];
  return n;
}

extern "C" int dhh_lanczos(dhh_problem* p, int nbrEigen, int blockSize, int maxIter, double precision,
                           double residual, uint64_t seed, double* eigenvalues, int* products, int* blocks,
                           double* seconds, double* d_vectors, int nbrVectors) {
  if (!p || !eigenvalues || nbrEigen <= 0) return DH_E_ARG;
  return guarded([&] {
    FullReorthogonalizedBlockLanczosAlgorithm lanczos(p->arch->arch.get(), p->hamiltonian.get(), nbrEigen, blockSize,
                                                      maxIter, precision, residual);
    lanczos.InitializeLanczosAlgorithm(seed);
    LanczosResult r = lanczos.Solve();
    for (int i = 0; i < nbrEigen; ++i) eigenvalues[i] = (i < (int)r.eigenvalues.size()) ? r.eigenvalues[i] : 0.0;
    if (products) *products = r.iterations;
    if (blocks) *blocks = r.blockIterations;
    if (seconds) *seconds = r.seconds;
    if (d_vectors && nbrVectors > 0) lanczos.GetEigenstates(nbrVectors, d_vectors);
    return r.converged ? DH_OK : DH_E_STATE;
  });
}

extern "C" int dhh_lanczos_disk(dhh_problem* p, const char* directory, int resume, int nbrEigen, int blockSize,
                                int maxIter, double precision, double residual, uint64_t seed, double* eigenvalues,
                                int* products, int* blocks, double* seconds, double* d_vectors, int nbrVectors) {
  if (!p || !directory || !eigenvalues || nbrEigen <= 0) return DH_E_ARG;
  return guarded([&] {
    FullReorthogonalizedBlockLanczosAlgorithm lanczos(p->arch->arch.get(), p->hamiltonian.get(), nbrEigen, blockSize,
                                                      maxIter, precision, residual);
    lanczos.SetDiskStorage(directory);
    if (resume)
      lanczos.ResumeLanczosAlgorithm();
    else
      lanczos.InitializeLanczosAlgorithm(seed);
    LanczosResult r = lanczos.Solve();
    for (int i = 0; i < nbrEigen; ++i) eigenvalues[i] = (i < (int)r.eigenvalues.size()) ? r.eigenvalues[i] : 0.0;
    if (products) *products = r.iterations;
    if (blocks) *blocks = r.blockIterations;
    if (seconds) *seconds = r.seconds;
    if (d_vectors && nbrVectors > 0) lanczos.GetEigenstates(nbrVectors, d_vectors);
    return r.converged ? DH_OK : DH_E_STATE;
  });
}

extern "C" int dhh_autotune(dhh_problem* p, const int* ks, int nbrKs, int repeats, double* table, int* bestK) {
  if (!p || !ks || nbrKs <= 0 || !table) return DH_E_ARG;
  return guarded([&] {
    GPUArchitecture* arch = p->arch->arch.get();
    dh_context* c = arch->Context();
    dh_hamiltonian* h = p->hamiltonian->Handle();
    const int64_t dim = p->hamiltonian->GetHilbertSpaceDimension();
    int64_t nnz = 0;
    arch->Check(dh_hamiltonian_count_elements(h, &nnz), "count elements");
    const bool spin = dynamic_cast<SpinChainHamiltonian*>(p->hamiltonian.get()) != nullptr;
    double bestRate = -1.0;
    int bestIdx = 0;
    std::vector<int> bestCfg(nbrKs * 4);
    for (int i = 0; i < nbrKs; ++i) {
      const int k = ks[i];
      double* x = arch->AllocateMemory(dim * k);
      double* y = arch->AllocateMemory(dim * k);
      arch->Check(dh_fill_random(c, x, dim * k, 11), "random");
      double bestMs = 1e300;
      int cfg[4] = {1, 256, DH_INDEX_RANK, 1};
      // both two-body row routines of the rank index explicitly (DH_INDEX_RANK picks one of them by
      // k); spin chains have one row routine per index
      const std::vector<int> modes = spin ? std::vector<int>{DH_INDEX_RANK, DH_INDEX_BSEARCH}
                                          : std::vector<int>{DH_INDEX_RANK_FREE, DH_INDEX_RANK_RANGE,
                                                             DH_INDEX_RANK_X4, DH_INDEX_BSEARCH};
      for (int mode : modes)
        for (int lpr : {1, 2, 4, 8}) {
          if (spin && lpr != 1) continue;
          for (int bs : {128, 256, 512, 1024}) {
            if (dh_hamiltonian_set_launch(h, lpr, bs, mode, 1) != DH_OK) continue;
            arch->Check(dh_hv(h, x, y, k, 0, 0, -1), "warm-up");
            arch->Check(dh_synchronize(c), "sync");
            auto t0 = std::chrono::steady_clock::now();
            for (int r = 0; r < repeats; ++r) arch->Check(dh_hv(h, x, y, k, 0, 0, -1), "timed product");
            arch->Check(dh_synchronize(c), "sync");
            double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() /
                        std::max(1, repeats);
            if (ms < bestMs) {
              bestMs = ms;
              cfg[0] = lpr;
              cfg[1] = bs;
              cfg[2] = mode;
            }
          }
        }
      arch->FreeMemory(x);
      arch->FreeMemory(y);
      double rate = (double)nnz * k / (bestMs * 1e-3);
      double* row = table + 6 * i;
      row[0] = k;
      row[1] = cfg[0];
      row[2] = cfg[1];
      row[3] = cfg[2];
      row[4] = bestMs;
      row[5] = rate;
      for (int q = 0; q < 4; ++q) bestCfg[4 * i + q] = cfg[q];
      if (rate > bestRate) {
        bestRate = rate;
        bestIdx = i;
      }
    }
    arch->Check(dh_hamiltonian_set_launch(h, bestCfg[4 * bestIdx], bestCfg[4 * bestIdx + 1],
                                          bestCfg[4 * bestIdx + 2], 1),
                "set launch");
    if (bestK) *bestK = ks[bestIdx];
    return DH_OK;
  });
}

extern "C" int dhh_sectors_solve(dhh_arch* a, int n, int lzMax, const int* totalLz, int nbrSectors,
                                 const double* pseudo, int nbrEigen, int blockSize, int maxIter, double precision,
                                 uint64_t seed, int batched, double memoryGB, double* eigenvalues, int* products,
                                 int* converged, double* seconds) {
  if (!a || !totalLz || !pseudo || !eigenvalues || nbrSectors <= 0 || nbrEigen <= 0) return DH_E_ARG;
  return guarded([&] {
    std::vector<int> lz(totalLz, totalLz + nbrSectors);
    std::vector<double> v(pseudo, pseudo + lzMax + 1);
    auto t0 = std::chrono::steady_clock::now();
    bool all = true;
    for (int i = 0; i < nbrSectors * nbrEigen; ++i) eigenvalues[i] = 0.0;
    if (batched) {
      MultiSectorBlockLanczos solver(a->arch.get(), n, lzMax, lz, v, nbrEigen, blockSize, maxIter, precision,
                                     (size_t)(memoryGB * 1073741824.0));
      solver.SetLaunch(g_sectorLaunch);
      solver.SetAloneRows(g_aloneRows);
      g_lastWaves = 0;
      std::vector<SectorResult> r = solver.Solve(seed);
      g_lastWaves = solver.NbrWaves();
      for (int s = 0; s < nbrSectors; ++s) {
        for (size_t e = 0; e < r[s].eigenvalues.size() && (int)e < nbrEigen; ++e)
          eigenvalues[(size_t)s * nbrEigen + e] = r[s].eigenvalues[e];
        if (products) products[s] = r[s].products;
        if (converged) converged[s] = r[s].converged;
        all = all && r[s].converged;
      }
    } else {
      g_lastWaves = 0;
      ProgressLog log;
      for (int s = 0; s < nbrSectors; ++s) {
        if (log.Due()) std::fprintf(stderr, "[sectors] sector %d of %d (2 Lz = %d)\n", s + 1, nbrSectors, lz[s]);
        FermionOnSphere space(a->arch.get(), n, lz[s], lzMax);
        ParticleOnSphereGenericHamiltonian h(&space, n, lzMax, v, a->arch.get());
        FullReorthogonalizedBlockLanczosAlgorithm lanczos(a->arch.get(), &h, nbrEigen, blockSize, maxIter, precision);
        lanczos.InitializeLanczosAlgorithm(seed);
        LanczosResult r = lanczos.Solve();
        for (size_t e = 0; e < r.eigenvalues.size() && (int)e < nbrEigen; ++e)
          eigenvalues[(size_t)s * nbrEigen + e] = r.eigenvalues[e];
        if (products) products[s] = r.iterations;
        if (converged) converged[s] = r.converged;
        all = all && r.converged;
      }
    }
    dh_synchronize(a->arch->Context());
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return all ? DH_OK : DH_E_STATE;
  });
}

extern "C" int dhh_sectors_autotune(dhh_arch* a, int n, int lzMax, const int* totalLz, int nbrSectors,
                                    const double* pseudo, int nbrEigen, const int* ks, int nbrKs, double fraction,
                                    int maxIter, uint64_t seed, double* table, int* best) {
  if (!a || !totalLz || !pseudo || !ks || !table || !best || nbrSectors <= 0 || nbrKs <= 0 || nbrEigen <= 0 ||
      !(fraction > 0.0))
    return DH_E_ARG;
  return guarded([&] {
    GPUArchitecture* arch = a->arch.get();
    dh_context* c = arch->Context();
    std::vector<double> v(pseudo, pseudo + lzMax + 1), F;
    std::vector<int> P;
    ParticleOnSphereGenericHamiltonian::EvaluateInteractionFactors(lzMax, v, F, P);
    // calibration sectors: every step-th of the sectors too large for the dense path, by size,
    // so that they hold about `fraction` of their rows and span their size range
    std::vector<std::pair<int64_t, int>> bySize;
    int64_t rowsAll = 0;
    for (int s = 0; s < nbrSectors; ++s) {
      const int64_t d = dh_fermion_sphere_dimension(n, totalLz[s], lzMax);
      if (d <= 0) throw std::invalid_argument("empty sector");
      rowsAll += d;
      if (d > 256) bySize.push_back({d, totalLz[s]});
    }
    if (bySize.empty()) throw std::invalid_argument("no sector large enough to tune on");
    std::sort(bySize.rbegin(), bySize.rend());
    const int step = std::max(1, (int)std::lround(1.0 / std::min(1.0, fraction)));
    std::vector<int> lz;
    int64_t rows = 0;
    for (size_t i = step / 2; i < bySize.size(); i += step) {
      lz.push_back(bySize[i].second);
      rows += bySize[i].first;
    }
    if (lz.empty()) {
      lz.push_back(bySize.back().second);
      rows = bySize.back().first;
    }
    SectorSolveResources set(arch);
    arch->Check(dh_sector_set_create(c, &set.set, n, lzMax, lz.data(), (int)lz.size()), "calibration set");
    arch->Check(dh_sector_set_factors(set.set, F.data(), P.data(), (int)F.size(), 0.0), "factors");
    const int lprs[] = {1, 2, 4}, blocks[] = {256, 1024}, modes[] = {DH_INDEX_RANK_RANGE, DH_INDEX_RANK_FREE};
    double bestSeconds = 1e300;
    int bestIdx = -1;
    for (int i = 0; i < nbrKs; ++i) {
      const int k = ks[i];
      double* row = table + 7 * i;
      for (int q = 0; q < 7; ++q) row[q] = 0.0;
      row[0] = k;
      double* x = nullptr;
      double* y = nullptr;
      try {
        x = arch->AllocateMemory(rows * k);
        y = arch->AllocateMemory(rows * k);
      } catch (const std::exception&) {
        if (x) arch->FreeMemory(x);
        row[5] = -1.0;  // the vectors of this block size do not fit
        continue;
      }
      arch->Check(dh_fill_random(c, x, rows * k, seed), "random block");
      double bestMs = 1e300;
      SectorLaunch shape;
      for (int lpr : lprs)
        for (int bs : blocks)
          for (int mode : modes) {
            arch->Check(dh_sector_set_launch(set.set, lpr, bs), "launch");
            arch->Check(dh_sector_set_row_mode(set.set, mode), "row mode");
            arch->Check(dh_sectors_hv(set.set, x, y, k, 0, nullptr), "warm-up product");
            arch->Check(dh_synchronize(c), "sync");
            auto t0 = std::chrono::steady_clock::now();
            for (int r = 0; r < 2; ++r) arch->Check(dh_sectors_hv(set.set, x, y, k, 0, nullptr), "timed product");
            arch->Check(dh_synchronize(c), "sync");
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / 2;
            if (ms < bestMs) {
              bestMs = ms;
              shape.lanesPerRow = lpr;
              shape.blockSize = bs;
              shape.rowMode = mode;
            }
          }
      arch->FreeMemory(x);
      arch->FreeMemory(y);
      // time to the eigenvalues of the calibration sectors with that shape
      MultiSectorBlockLanczos solver(arch, n, lzMax, lz, v, nbrEigen, k, maxIter, 1e-14, 0);
      solver.SetLaunch(shape);
      solver.SetAloneRows(g_aloneRows);
      double seconds = 0.0;
      std::vector<SectorResult> res = solver.Solve(seed, &seconds);
      int products = 0;
      for (const SectorResult& r : res) products += r.products;
      row[1] = shape.lanesPerRow;
      row[2] = shape.blockSize;
      row[3] = shape.rowMode;
      row[4] = bestMs;
      row[5] = seconds;
      row[6] = products;
      if (seconds < bestSeconds) {
        bestSeconds = seconds;
        bestIdx = i;
      }
    }
    if (bestIdx < 0) throw std::runtime_error("no block size fits the device");
    for (int q = 0; q < 4; ++q) best[q] = (int)table[7 * bestIdx + q];
    return DH_OK;
  });
}

extern "C" int dhh_lattice_sectors_solve(dhh_arch* a, int n, int nx, int ny, double t1, double t2, double tpp,
                                         double mus, double uPot, double vPot, int flat, const int* kx, const int* ky,
                                         int nbrSectors, int nbrEigen, int blockSize, int maxIter, double precision,
                                         uint64_t seed, double memoryGB, double* eigenvalues, int* products,
                                         int* converged, double* seconds) {
  if (!a || !kx || !ky || !eigenvalues || nbrSectors <= 0 || nbrEigen <= 0) return DH_E_ARG;
  return guarded([&] {
    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::complex<double>> u;
    std::vector<double> e, f;
    std::vector<int> cnt, pairs;
    CheckerboardLowerBand(nx, ny, t1, t2, tpp, mus, 0.0, 0.0, u, e);
    CheckerboardSingleBandFactors(nx, ny, u, uPot, vPot, flat != 0, cnt, pairs, f);
    LatticeSectorsBlockLanczos solver(a->arch.get(), n, nx, ny, std::vector<int>(kx, kx + nbrSectors),
                                      std::vector<int>(ky, ky + nbrSectors), cnt, pairs, f, nbrEigen, blockSize,
                                      maxIter, precision, (size_t)(memoryGB * 1073741824.0));
    SectorLaunch launch = g_sectorLaunch;
    launch.rowMode = DH_INDEX_RANK;  // momentum sets have one row routine
    solver.SetLaunch(launch);
    g_lastWaves = 0;
    std::vector<SectorResult> r = solver.Solve(seed);
    g_lastWaves = solver.NbrWaves();
    bool all = true;
    for (int s = 0; s < nbrSectors; ++s) {
      for (int q = 0; q < nbrEigen; ++q)
        eigenvalues[(size_t)s * nbrEigen + q] = q < (int)r[s].eigenvalues.size() ? r[s].eigenvalues[q] : 0.0;
      if (products) products[s] = r[s].products;
      if (converged) converged[s] = r[s].converged;
      all = all && r[s].converged;
    }
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return all ? DH_OK : DH_E_STATE;
  });
}

extern "C" int dhh_checkerboard_lower_band(int nx, int ny, double t1, double t2, double tpp, double mus, double gx,
                                           double gy, double* u, double* energies) {
  if (!u || !energies || nx <= 0 || ny <= 0) return DH_E_ARG;
  std::vector<std::complex<double>> uu;
  std::vector<double> e;
  CheckerboardLowerBand(nx, ny, t1, t2, tpp, mus, gx, gy, uu, e);
  for (size_t i = 0; i < uu.size(); ++i) {
    u[2 * i] = uu[i].real();
    u[2 * i + 1] = uu[i].imag();
  }
  std::memcpy(energies, e.data(), sizeof(double) * e.size());
  return DH_OK;
}

extern "C" int dhh_checkerboard_factors(int nx, int ny, double t1, double t2, double tpp, double mus, double uPot,
                                        double vPot, int flat, int* sumPairCount, int* sumPairs, int* nbrPairs,
                                        double* factors, int64_t* nbrFactors) {
  if (!sumPairCount || !sumPairs || !nbrPairs || !factors || !nbrFactors) return DH_E_ARG;
  std::vector<std::complex<double>> u;
  std::vector<double> e, f;
  std::vector<int> cnt, pairs;
  CheckerboardLowerBand(nx, ny, t1, t2, tpp, mus, 0.0, 0.0, u, e);
  CheckerboardSingleBandFactors(nx, ny, u, uPot, vPot, flat != 0, cnt, pairs, f);
  if ((int)pairs.size() / 2 > *nbrPairs || (int64_t)f.size() / 2 > *nbrFactors) return DH_E_RANGE;
  std::memcpy(sumPairCount, cnt.data(), sizeof(int) * cnt.size());
  std::memcpy(sumPairs, pairs.data(), sizeof(int) * pairs.size());
  std::memcpy(factors, f.data(), sizeof(double) * f.size());
  *nbrPairs = (int)pairs.size() / 2;
  *nbrFactors = (int64_t)f.size() / 2;
  return DH_OK;
}

extern "C" int dhh_torus_factors(int nbrFlux, double ratio, const double* pseudo, int nbrPseudo, int coulombLevel,
                                 int* sumPairCount, int* sumPairs, int* nbrPairs, double* factors,
                                 int64_t* nbrFactors) {
  if (!sumPairCount || !sumPairs || !nbrPairs || !factors || !nbrFactors || nbrPseudo < 0 || (nbrPseudo && !pseudo))
    return DH_E_ARG;
  return guarded([&] {
    std::vector<int> cnt, pairs;
    std::vector<double> f;
    TorusInteractionFactors(nbrFlux, ratio, std::vector<double>(pseudo, pseudo + nbrPseudo), coulombLevel, cnt,
                            pairs, f);
    if ((int)pairs.size() / 2 > *nbrPairs || (int64_t)f.size() / 2 > *nbrFactors) return (int)DH_E_RANGE;
    std::memcpy(sumPairCount, cnt.data(), sizeof(int) * cnt.size());
    std::memcpy(sumPairs, pairs.data(), sizeof(int) * pairs.size());
    std::memcpy(factors, f.data(), sizeof(double) * f.size());
    *nbrPairs = (int)pairs.size() / 2;
    *nbrFactors = (int64_t)f.size() / 2;
    return (int)DH_OK;
  });
}

extern "C" int dhh_torus_sectors_solve(dhh_arch* a, int n, int nbrFlux, double ratio, const double* pseudo,
                                       int nbrPseudo, int coulombLevel, const int* ky, int nbrSectors, int nbrEigen,
                                       int blockSize, int maxIter, double precision, uint64_t seed, double memoryGB,
                                       double* eigenvalues, int* products, int* converged, double* seconds) {
  if (!a || !ky || !eigenvalues || nbrSectors <= 0 || nbrEigen <= 0 || nbrPseudo < 0 || (nbrPseudo && !pseudo))
    return DH_E_ARG;
  return guarded([&] {
    auto t0 = std::chrono::steady_clock::now();
    std::vector<int> cnt, pairs;
    std::vector<double> f;
    TorusInteractionFactors(nbrFlux, ratio, std::vector<double>(pseudo, pseudo + nbrPseudo), coulombLevel, cnt,
                            pairs, f);
    for (int s = 0; s < nbrSectors; ++s)
      if (ky[s] < 0 || ky[s] >= nbrFlux) throw std::invalid_argument("Ky outside [0, nbr_flux)");
    LatticeSectorsBlockLanczos solver(a->arch.get(), n, 1, nbrFlux, std::vector<int>(nbrSectors, 0),
                                      std::vector<int>(ky, ky + nbrSectors), cnt, pairs, f, nbrEigen, blockSize,
                                      maxIter, precision, (size_t)(memoryGB * 1073741824.0));
    SectorLaunch launch = g_sectorLaunch;
    launch.rowMode = DH_INDEX_RANK;  // momentum sets have one row routine
    solver.SetLaunch(launch);
    g_lastWaves = 0;
    std::vector<SectorResult> r = solver.Solve(seed);
    g_lastWaves = solver.NbrWaves();
    bool all = true;
    for (int s = 0; s < nbrSectors; ++s) {
      for (int q = 0; q < nbrEigen; ++q)
        eigenvalues[(size_t)s * nbrEigen + q] = q < (int)r[s].eigenvalues.size() ? r[s].eigenvalues[q] : 0.0;
      if (products) products[s] = r[s].products;
      if (converged) converged[s] = r[s].converged;
      all = all && r[s].converged;
    }
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return all ? (int)DH_OK : (int)DH_E_STATE;
  });
}

extern "C" int dhh_torus_translation_sectors_solve(dhh_arch* a, int n, int nbrFlux, double ratio,
                                                   const double* pseudo, int nbrPseudo, int coulombLevel,
                                                   const int* kx, const int* ky, int nbrSectors, int nbrEigen,
                                                   int blockSize, int maxIter, double precision, uint64_t seed,
                                                   double memoryGB, double* eigenvalues, int* products,
                                                   int* converged, double* seconds) {
  if (!a || !kx || !ky || !eigenvalues || nbrSectors <= 0 || nbrEigen <= 0 || nbrPseudo < 0 || (nbrPseudo && !pseudo))
    return DH_E_ARG;
  return guarded([&] {
    auto t0 = std::chrono::steady_clock::now();
    std::vector<int> cnt, pairs;
    std::vector<double> f;
    TorusInteractionFactors(nbrFlux, ratio, std::vector<double>(pseudo, pseudo + nbrPseudo), coulombLevel, cnt,
                            pairs, f);
    LatticeSectorsBlockLanczos solver(a->arch.get(), n, 1, nbrFlux, std::vector<int>(kx, kx + nbrSectors),
                                      std::vector<int>(ky, ky + nbrSectors), cnt, pairs, f, nbrEigen, blockSize,
                                      maxIter, precision, (size_t)(memoryGB * 1073741824.0));
    solver.UseTorusTranslations();
    SectorLaunch launch = g_sectorLaunch;
    launch.rowMode = DH_INDEX_RANK;  // momentum sets have one row routine
    solver.SetLaunch(launch);
    g_lastWaves = 0;
    std::vector<SectorResult> r = solver.Solve(seed);
    g_lastWaves = solver.NbrWaves();
    bool all = true;
    for (int s = 0; s < nbrSectors; ++s) {
      for (int q = 0; q < nbrEigen; ++q)
        eigenvalues[(size_t)s * nbrEigen + q] = q < (int)r[s].eigenvalues.size() ? r[s].eigenvalues[q] : 0.0;
      if (products) products[s] = r[s].products;
      if (converged) converged[s] = r[s].converged;
      all = all && r[s].converged;
    }
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return all ? (int)DH_OK : (int)DH_E_STATE;
  });
}

extern "C" int dhh_coulomb_pseudopotentials(int nbrFlux, int round, double* out) {
  if (!out || nbrFlux < 0) return DH_E_ARG;
  std::vector<double> v = EvaluatePseudopotentials(nbrFlux);
  if (round) v = RoundToFileDigits(v);
  std::memcpy(out, v.data(), sizeof(double) * v.size());
  return DH_OK;
}

extern "C" double dhh_clebsch_gordan(int j1, int j2, int m1, int m2, int j) {
  return ClebschGordanCoefficients::Coefficient(j1, j2, m1, m2, j);
}

extern "C" int dhh_sphere_factors(int lzMax, const double* pseudo, double* factors, int* pairOffset, int* nbr) {
  if (!pseudo || !factors || !pairOffset || !nbr) return DH_E_ARG;
  std::vector<double> v(pseudo, pseudo + lzMax + 1), f;
  std::vector<int> off;
  ParticleOnSphereGenericHamiltonian::EvaluateInteractionFactors(lzMax, v, f, off);
  if ((int)f.size() > *nbr) return DH_E_RANGE;
  std::memcpy(factors, f.data(), sizeof(double) * f.size());
  std::memcpy(pairOffset, off.data(), sizeof(int) * off.size());
  *nbr = (int)f.size();
  return DH_OK;
}

extern "C" int dhh_sphere_channel_table(int lzMax, int* momenta, double* table, int capacity) {
  std::vector<int> m;
  std::vector<double> t;
  ParticleOnSphereGenericHamiltonian::EvaluateChannelTable(lzMax, m, t);
  if ((int)t.size() > capacity) return DH_E_RANGE;
  std::memcpy(momenta, m.data(), sizeof(int) * m.size());
  std::memcpy(table, t.data(), sizeof(double) * t.size());
  return (int)m.size();
}

extern "C" int dhh_read_pseudopotentials(const char* file, int lzMax, double* out) {
  if (!file || !out) return DH_E_ARG;
  std::vector<double> v;
  if (!FQHESphereGetPseudopotentials(file, lzMax, v)) return DH_E_ARG;
  std::memcpy(out, v.data(), sizeof(double) * v.size());
  return DH_OK;
}

extern "C" int dhh_band_eigenvalues(int n, int b, const double* a, double* ev) {
  if (!a || !ev || n <= 0 || b < 1) return DH_E_ARG;
  return guarded([&] {
    std::vector<double> m(a, a + (size_t)n * n), d, e, w;
    BandToTridiagonal(n, b, m, d, e, nullptr);
    TridiagonalEigen(n, d, e, w, nullptr);
    std::memcpy(ev, w.data(), sizeof(double) * n);
    return DH_OK;
  });
}
