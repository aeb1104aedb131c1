// All Lz sectors of a fermions-on-the-sphere problem, solved together.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "diagham_host.hpp"

namespace dhh {

namespace {
bool cholesky_upper_k(int k, const double* G, std::vector<double>& R) {
  R.assign((size_t)k * k, 0.0);
  for (int i = 0; i < k; ++i) {
    double s = G[(size_t)i * k + i];
    for (int p = 0; p < i; ++p) s -= R[(size_t)p * k + i] * R[(size_t)p * k + i];
    if (!(s > 1e-28 * std::max(1.0, std::fabs(G[(size_t)i * k + i])))) return false;
    double rii = std::sqrt(s);
    R[(size_t)i * k + i] = rii;
    for (int j = i + 1; j < k; ++j) {
      double t = G[(size_t)i * k + j];
      for (int p = 0; p < i; ++p) t -= R[(size_t)p * k + i] * R[(size_t)p * k + j];
      R[(size_t)i * k + j] = t / rii;
    }
  }
  return true;
}

std::vector<double> upper_inv(int k, const std::vector<double>& R) {
  std::vector<double> X((size_t)k * k, 0.0);
  for (int j = 0; j < k; ++j) {
    X[(size_t)j * k + j] = 1.0 / R[(size_t)j * k + j];
    for (int i = j - 1; i >= 0; --i) {
      double s = 0.0;
      for (int p = i + 1; p <= j; ++p) s += R[(size_t)i * k + p] * X[(size_t)p * k + j];
      X[(size_t)i * k + j] = -s / R[(size_t)i * k + i];
    }
  }
  return X;
}

std::vector<double> mul(int k, const std::vector<double>& A, const std::vector<double>& B) {
  std::vector<double> C((size_t)k * k, 0.0);
  for (int i = 0; i < k; ++i)
    for (int p = 0; p < k; ++p)
      for (int j = 0; j < k; ++j) C[(size_t)i * k + j] += A[(size_t)i * k + p] * B[(size_t)p * k + j];
  return C;
}
}  // namespace

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
      MemoryBytes(memoryBytes ? memoryBytes : (size_t)160 << 30),
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
  // the rest in waves whose Krylov storage fits the memory budget
  const int k = BlockSize;
  const int maxBlocks = std::max(3, MaxIter / k + 1);
  std::vector<int> wave;
  size_t used = 0;
  Waves = 0;
  for (int s : lanczosSectors) {
    size_t need = (size_t)out[s].dimension * k * (maxBlocks + 2) * sizeof(double);
    if (!wave.empty() && used + need > MemoryBytes) {
      SolveWave(wave, seed, out);
      ++Waves;
      wave.clear();
      used = 0;
    }
    wave.push_back(s);
    used += need;
  }
  if (!wave.empty()) {
    SolveWave(wave, seed, out);
    ++Waves;
  }
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return out;
}

void MultiSectorBlockLanczos::SolveWave(const std::vector<int>& sectors, uint64_t seed,
                                        std::vector<SectorResult>& out) {
  dh_context* c = Arch->Context();
  const int k = BlockSize, kk = k * k;
  const int na = (int)sectors.size();
  std::vector<int> lz;
  for (int s : sectors) lz.push_back(TotalLz[s]);
  dh_sector_set* S = nullptr;
  Arch->Check(dh_sector_set_create(c, &S, NbrFermions, LzMax, lz.data(), na), "sector set");
  Arch->Check(dh_sector_set_factors(S, Factors.data(), PairOffset.data(), (int)Factors.size(), 0.0), "factors");
  std::vector<int64_t> off(na + 1);
  dh_sector_set_offsets(S, off.data());
  const int64_t total = off.back();
  Arch->Check(dh_sector_set_launch(S, total < 4ll * 256 * 256 ? 4 : 1, 256), "launch");
  int maxBlocks = 2;
  std::vector<int> blockCap(na);
  for (int i = 0; i < na; ++i) {
    int64_t d = off[i + 1] - off[i];
    blockCap[i] = (int)std::min<int64_t>(MaxIter / k + 1, (d + k - 1) / k + 1);
    maxBlocks = std::max(maxBlocks, blockCap[i]);
  }
  double* basis = Arch->AllocateMemory((int64_t)(maxBlocks + 1) * total * k);
  auto blockPtr = [&](int j) { return basis + (int64_t)j * total * k; };
  Arch->Check(dh_fill_random(c, basis, total * k, seed), "random start");

  std::vector<unsigned char> active(na, 1);
  std::vector<std::vector<std::vector<double>>> A(na), B(na);
  std::vector<std::vector<double>> prev(na);
  std::vector<int> lastCheck(na, -1);
  std::vector<char> done(na, 0), exhausted(na, 0);

  // segmented Cholesky-QR (two passes) of block W over the active sectors; returns R per
  // active sector (empty on rank deficiency)
  auto orthonormalize = [&](double* W, std::vector<std::vector<double>>& R) {
    std::vector<int> act;
    for (int i = 0; i < na; ++i)
      if (active[i]) act.push_back(i);
    R.assign(na, std::vector<double>());
    for (int i : act) {
      R[i].assign((size_t)kk, 0.0);
      for (int p = 0; p < k; ++p) R[i][(size_t)p * k + p] = 1.0;
    }
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<double> G((size_t)act.size() * kk);
      Arch->Check(dh_sectors_block_dots(S, W, 1, W, k, (const unsigned char*)active.data(), G.data()), "gram");
      std::vector<double> M((size_t)act.size() * kk, 0.0);
      for (size_t a = 0; a < act.size(); ++a) {
        std::vector<double> Rp;
        int i = act[a];
        if (!R[i].empty() && cholesky_upper_k(k, &G[a * kk], Rp)) {
          std::vector<double> inv = upper_inv(k, Rp);
          std::memcpy(&M[a * kk], inv.data(), sizeof(double) * kk);
          R[i] = mul(k, Rp, R[i]);
        } else {
          R[i].clear();
          for (int p = 0; p < k; ++p) M[a * kk + p * k + p] = 1.0;
        }
      }
      Arch->Check(dh_sectors_block_rmul(S, W, k, (const unsigned char*)active.data(), M.data()), "rmul");
    }
  };

  {
    std::vector<std::vector<double>> R;
    orthonormalize(blockPtr(0), R);
    for (int i = 0; i < na; ++i)
      if (R[i].empty()) throw std::runtime_error("random start block is rank deficient");
  }
  for (int j = 0;; ++j) {
    std::vector<int> act;
    for (int i = 0; i < na; ++i)
      if (!done[i]) act.push_back(i);
    for (int i = 0; i < na; ++i) active[i] = done[i] ? 0 : 1;
    if (act.empty() || j + 1 > maxBlocks) break;
    const unsigned char* am = (const unsigned char*)active.data();
    double* X = blockPtr(j);
    double* W = blockPtr(j + 1);
    Arch->Check(dh_sectors_hv(S, X, W, k, 0, am), "sector products");
    const size_t per = (size_t)(j + 1) * kk;
    std::vector<double> C(act.size() * per);
    std::vector<std::vector<double>> Aj(na, std::vector<double>((size_t)kk, 0.0));
    for (int pass = 0; pass < 2; ++pass) {
      Arch->Check(dh_sectors_block_dots(S, basis, j + 1, W, k, am, C.data()), "reorth dots");
      for (size_t a = 0; a < act.size(); ++a)
        for (int q = 0; q < kk; ++q) Aj[act[a]][q] += C[a * per + (size_t)j * kk + q];
      for (auto& v : C) v = -v;
      Arch->Check(dh_sectors_block_axpy(S, basis, j + 1, W, k, am, C.data()), "reorth axpy");
    }
    std::vector<std::vector<double>> R;
    orthonormalize(W, R);
    for (int i : act) {
      auto& a = Aj[i];
      for (int p = 0; p < k; ++p)
        for (int q = p + 1; q < k; ++q) a[(size_t)p * k + q] = a[(size_t)q * k + p] = 0.5 * (a[(size_t)p * k + q] + a[(size_t)q * k + p]);
      A[i].push_back(a);
      if (R[i].empty()) {
        exhausted[i] = 1;  // Krylov space spans the sector (or an invariant subspace of it)
        B[i].push_back(std::vector<double>((size_t)kk, 0.0));
      } else {
        B[i].push_back(R[i]);
      }
      out[sectors[i]].products += k;
    }
    // convergence per sector (strong test of FullReorthogonalizedLanczosAlgorithm.cc:325)
#pragma omp parallel for schedule(dynamic, 1)
    for (size_t a = 0; a < act.size(); ++a) {
      int i = act[a];
      int m = (int)A[i].size();
      bool last = exhausted[i] || m >= blockCap[i];
      if (!last && !(k == 1 || lastCheck[i] < 0 || m - lastCheck[i] >= std::max(1, m / 10))) continue;
      if (!last && m * k < NbrEigenvalue + 2) continue;
      lastCheck[i] = m;
      std::vector<double> ev;
      BlockTridiagonalEigen(A[i], B[i], k, m, ev, nullptr);
      int want = std::min<int>(NbrEigenvalue, (int)ev.size());
      std::vector<double> cur(ev.begin(), ev.begin() + want);
      bool ok = !prev[i].empty() && (int)ev.size() > NbrEigenvalue && (int)prev[i].size() == want;
      if (ok)
        for (int e = 0; e < want; ++e)
          if (std::fabs(cur[e] - prev[i][e]) > Precision * std::fabs(cur[e])) {
            ok = false;
            break;
          }
      prev[i] = cur;
      SectorResult& r = out[sectors[i]];
      r.eigenvalues = cur;
      if (ok || (exhausted[i] && (int64_t)m * k >= off[i + 1] - off[i])) {
        r.converged = true;
        done[i] = 1;
      } else if (last) {
        done[i] = 1;
      }
    }
  }
  Arch->Check(dh_synchronize(c), "sync");
  Arch->FreeMemory(basis);
  dh_sector_set_destroy(S);
}

}  // namespace dhh
