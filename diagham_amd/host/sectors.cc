// All Lz sectors of a fermions-on-the-sphere problem, solved together.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>

#include "diagham_host.hpp"

namespace dhh {

// block operations over the rows of a sector set (real or complex vectors: kr doubles per row)
SectorSetOps::SectorSetOps(GPUArchitecture* arch, dh_sector_set* set, int k, int kr, int64_t rows)
    : Arch(arch), Set(set), K(k), KR(kr), Rows(rows) {}

void SectorSetOps::Product(const double* X, double* W, const unsigned char* active) {
  Arch->Check(dh_sectors_hv(Set, X, W, K, 0, active), "sector products");
}
void SectorSetOps::Dots(const double* V, int nb, const double* W, const unsigned char* active, double* out) {
  Arch->Check(dh_sectors_block_dots(Set, V, nb, W, KR, active, out), "sector dots");
}
void SectorSetOps::Axpy(const double* V, int nb, double* W, const unsigned char* active, const double* C) {
  Arch->Check(dh_sectors_block_axpy(Set, V, nb, W, KR, active, C), "sector axpy");
}
void SectorSetOps::Rmul(double* W, const unsigned char* active, const double* M) {
  Arch->Check(dh_sectors_block_rmul(Set, W, KR, active, M), "sector rmul");
}
void SectorSetOps::Random(double* W, uint64_t seed) {
  Arch->Check(dh_fill_random(Arch->Context(), W, Rows * KR, seed), "random block");
}

MultiSectorBlockLanczos::MultiSectorBlockLanczos(GPUArchitecture* architecture, int nbrFermions, int lzMax,
                                                 const std::vector<int>& totalLz,
                                                 const std::vector<double>& pseudoPotential, int nbrEigenvalue,
                                                 int blockSize, int maxIter, double precision, size_t memoryBytes)
    : Arch(architecture),
      NbrFermions(nbrFermions),
      LzMax(lzMax),
      NbrEigenvalue(nbrEigenvalue),
      BlockSize(blockSize),
      MaxIter(maxIter),
      Precision(precision),
      MemoryBytes(memoryBytes ? memoryBytes : SIZE_MAX),
      TotalLz(totalLz),
      Pseudo(pseudoPotential) {
  if (BlockSize != 1 && BlockSize != 2 && BlockSize != 4 && BlockSize != 8 && BlockSize != 16)
    throw std::invalid_argument("block size must be 1, 2, 4, 8 or 16");
  ParticleOnSphereGenericHamiltonian::EvaluateInteractionFactors(lzMax, pseudoPotential, Factors, PairOffset);
}

std::vector<SectorResult> MultiSectorBlockLanczos::Solve(uint64_t seed, double* seconds) {
  auto t0 = std::chrono::steady_clock::now();
  const int ns = (int)TotalLz.size();
  std::vector<SectorResult> out(ns);
  std::vector<int> lanczosSectors, denseSectors;
  for (int s = 0; s < ns; ++s) {
    out[s].totalLz = TotalLz[s];
    out[s].dimension = dh_fermion_sphere_dimension(NbrFermions, TotalLz[s], LzMax);
    if (out[s].dimension <= 0) throw std::invalid_argument("empty sector");
    if (out[s].dimension <= std::max<int64_t>(64, 4 * BlockSize))
      denseSectors.push_back(s);
    else
      lanczosSectors.push_back(s);
  }
  dh_context* c = Arch->Context();
  // small sectors: dense H from unit-vector products over all of them at once
  if (!denseSectors.empty()) {
    std::vector<int> lz;
    for (int s : denseSectors) lz.push_back(TotalLz[s]);
    dh_sector_set* S = nullptr;
    Arch->Check(dh_sector_set_create(c, &S, NbrFermions, LzMax, lz.data(), (int)lz.size()), "dense sector set");
    Arch->Check(dh_sector_set_factors(S, Factors.data(), PairOffset.data(), (int)Factors.size(), 0.0), "factors");
    std::vector<int64_t> off(lz.size() + 1);
    dh_sector_set_offsets(S, off.data());
    int64_t total = off.back();
    int64_t maxD = 0;
    for (size_t i = 0; i < lz.size(); ++i) maxD = std::max(maxD, off[i + 1] - off[i]);
    double* x = Arch->AllocateMemory(total);
    double* y = Arch->AllocateMemory(total);
    std::vector<std::vector<double>> H(lz.size());
    for (size_t i = 0; i < lz.size(); ++i) H[i].assign((size_t)(off[i + 1] - off[i]) * (off[i + 1] - off[i]), 0.0);
    std::vector<double> hx(total), hy(total);
    for (int64_t col = 0; col < maxD; ++col) {
      std::fill(hx.begin(), hx.end(), 0.0);
      for (size_t i = 0; i < lz.size(); ++i)
        if (col < off[i + 1] - off[i]) hx[off[i] + col] = 1.0;
      Arch->Check(dh_memcpy_h2d(c, x, hx.data(), sizeof(double) * total), "unit vectors");
      Arch->Check(dh_sectors_hv(S, x, y, 1, 0, nullptr), "dense products");
      Arch->Check(dh_memcpy_d2h(c, hy.data(), y, sizeof(double) * total), "columns");
      for (size_t i = 0; i < lz.size(); ++i) {
        int64_t d = off[i + 1] - off[i];
        if (col < d)
          for (int64_t r = 0; r < d; ++r) H[i][(size_t)r * d + col] = hy[off[i] + r];
      }
    }
    Arch->FreeMemory(x);
    Arch->FreeMemory(y);
    dh_sector_set_destroy(S);
    for (size_t i = 0; i < lz.size(); ++i) {
      int n = (int)(off[i + 1] - off[i]);
      std::vector<double>& h = H[i];
      for (int r = 0; r < n; ++r)
        for (int q = r + 1; q < n; ++q) h[(size_t)r * n + q] = h[(size_t)q * n + r] = 0.5 * (h[(size_t)r * n + q] + h[(size_t)q * n + r]);
      std::vector<double> d, e, w;
      BandToTridiagonal(n, std::max(1, n - 1), h, d, e, nullptr);
      TridiagonalEigen(n, d, e, w, nullptr);
      SectorResult& r = out[denseSectors[i]];
      r.eigenvalues.assign(w.begin(), w.begin() + std::min(NbrEigenvalue, n));
      r.products = (int)maxD;
      r.converged = true;
    }
  }
  // sectors of at least AloneRows rows one at a time through the single-space solver (each fills
  // the device alone), largest first
  std::vector<int> batchSectors, aloneSectors;
  for (int s : lanczosSectors) (AloneRows > 0 && out[s].dimension >= AloneRows ? aloneSectors : batchSectors).push_back(s);
  std::stable_sort(aloneSectors.begin(), aloneSectors.end(),
                   [&](int a, int b) { return out[a].dimension > out[b].dimension; });
  ProgressLog log;
  for (size_t i = 0; i < aloneSectors.size(); ++i) {
    const int s = aloneSectors[i];
    if (log.Due())
      std::fprintf(stderr, "[sectors] alone: sector %zu of %zu (2 Lz = %d)\n", i + 1, aloneSectors.size(), TotalLz[s]);
    FermionOnSphere space(Arch, NbrFermions, TotalLz[s], LzMax);
    ParticleOnSphereGenericHamiltonian h(&space, NbrFermions, LzMax, Pseudo, Arch);
    FullReorthogonalizedBlockLanczosAlgorithm lanczos(Arch, &h, NbrEigenvalue, BlockSize, MaxIter, Precision);
    lanczos.InitializeLanczosAlgorithm(seed);
    LanczosResult r = lanczos.Solve();
    out[s].eigenvalues = r.eigenvalues;
    out[s].products = r.iterations;
    out[s].blocks = r.blockIterations;
    out[s].converged = r.converged;
  }
  lanczosSectors = batchSectors;
  // the rest in waves whose expected Krylov storage fits the device (RunSectorWaves)
  std::vector<int64_t> dims;
  for (int s : lanczosSectors) dims.push_back(out[s].dimension);
  std::vector<SectorResult> part(lanczosSectors.size());
  Waves = RunSectorWaves(
      dims, BlockSize, BlockSize, NbrEigenvalue, MaxIter, MemoryBytes, part,
      [&](const std::vector<int>& wave, int blocks) {
        std::vector<int> sectors;
        for (int i : wave) sectors.push_back(lanczosSectors[i]);
        SolveWave(sectors, seed, out, blocks);
        for (int i : wave) part[i] = out[lanczosSectors[i]];
      },
      [&] { return DeviceFreeBytes(Arch); });
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return out;
}

void MultiSectorBlockLanczos::SolveWave(const std::vector<int>& sectors, uint64_t seed,
                                        std::vector<SectorResult>& out, int expectedBlocks) {
  dh_context* c = Arch->Context();
  const int k = BlockSize;
  const int na = (int)sectors.size();
  std::vector<int> lz;
  for (int s : sectors) lz.push_back(TotalLz[s]);
  SectorSolveResources res(Arch);
  Arch->Check(dh_sector_set_create(c, &res.set, NbrFermions, LzMax, lz.data(), na), "sector set");
  dh_sector_set* S = res.set;
  Arch->Check(dh_sector_set_factors(S, Factors.data(), PairOffset.data(), (int)Factors.size(), 0.0), "factors");
  std::vector<int64_t> off(na + 1);
  dh_sector_set_offsets(S, off.data());
  const int64_t total = off.back();
  // small sets: more lanes per row for enough wavefronts; large ones: the fastest shapes of the
  // all-sector product of N=12 2S=33 measured on MI355X (k >= 4: two lanes per row, 1024-thread
  // workgroups; k < 4: one lane per row, 1024-thread workgroups, 236 ms against 416 ms at 256
  // threads: the workgroup's tables, with the 16-bit low rank table, take ~61 KiB of LDS)
  if (Launch.lanesPerRow > 0) {
    Arch->Check(dh_sector_set_launch(S, Launch.lanesPerRow, Launch.blockSize), "launch");
    Arch->Check(dh_sector_set_row_mode(S, Launch.rowMode), "row mode");
  } else if (total < 4ll * 256 * 256)
    Arch->Check(dh_sector_set_launch(S, 4, 512), "launch");
  else if (BlockSize >= 4)
    Arch->Check(dh_sector_set_launch(S, 2, 1024), "launch");
  else
    Arch->Check(dh_sector_set_launch(S, 1, 1024), "launch");
  std::vector<int64_t> dims(na);
  for (int i = 0; i < na; ++i) dims[i] = off[i + 1] - off[i];
  std::vector<int> blockCap;
  const int maxBlocks = SectorBlockCaps(dims, k, MaxIter, blockCap);
  // the blocks the planner expects the wave to need, or the usual start when it expects fewer
  const int capacity = std::min(maxBlocks, std::max(InitialKrylovBlocks(k, maxBlocks), expectedBlocks));
  res.basis.push_back(Arch->AllocateMemory((int64_t)(capacity + 2) * total * k));
  SectorSetOps ops(Arch, S, k, k, total);
  SegmentedBlockLanczos<double> lanczos(&ops, dims, k, maxBlocks, res.basis[0], total * k);
  lanczos.SetGrowth(capacity, [&](int& blocks, int need) {
    return GrowKrylovStorage(Arch, res.basis, blocks, need, total * k);
  });
  std::vector<SectorResult*> results;
  for (int s : sectors) results.push_back(&out[s]);
  RunSectorLanczos(lanczos, blockCap, NbrEigenvalue, Precision, seed, results);
  Arch->Check(dh_synchronize(c), "sync");
}

}  // namespace dhh
