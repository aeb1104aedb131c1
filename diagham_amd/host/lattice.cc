// Chern-band lattice model in momentum space: one-body band, projected two-body factors and
// a complex block Lanczos over many momentum sectors at once.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <stdexcept>

#include "diagham_host.hpp"

namespace dhh {

using cplx = std::complex<double>;

// TightBindingModelCheckerboardLattice (FTI/src/Tools/FTITightBinding/TightBindingModelCheckerboardLattice.cc:206,
// the form used with --single-band): H(k) = [[d1 + d3, B1], [conj(B1), d1 - d3]] with
// B1 = 4 t1 (cos(Kx/2) cos(Ky/2) cos(pi/4) + i sin(Kx/2) sin(Ky/2) sin(pi/4)), d1 = 4 t'' cos Kx cos Ky,
// d3 = mu_s + 2 t2 (cos Kx - cos Ky), K = 2 pi (k + gamma) / N.  Lower-band eigenvector v(k) in closed
// form, gauge v_0 >= 0 (v_1 >= 0 if v_0 = 0); returned as OneBodyBasis[k][0][alpha] = conj(v_alpha(k)).
void CheckerboardLowerBand(int nx, int ny, double t1, double t2, double tpp, double mus, double gx, double gy,
                           std::vector<cplx>& u, std::vector<double>& energies) {
  u.assign((size_t)nx * ny * 2, 0.0);
  energies.assign((size_t)nx * ny * 2, 0.0);
  for (int kx = 0; kx < nx; ++kx)
    for (int ky = 0; ky < ny; ++ky) {
      const int index = kx * ny + ky;
      const double Kx = 2.0 * M_PI * (kx + gx) / nx, Ky = 2.0 * M_PI * (ky + gy) / ny;
      const cplx B1 = 4.0 * t1 * cplx(std::cos(0.5 * Kx) * std::cos(0.5 * Ky) * std::cos(0.25 * M_PI),
                                      std::sin(0.5 * Kx) * std::sin(0.5 * Ky) * std::sin(0.25 * M_PI));
      const double d1 = 4.0 * tpp * std::cos(Kx) * std::cos(Ky);
      const double d3 = mus + 2.0 * t2 * (std::cos(Kx) - std::cos(Ky));
      const double r = std::sqrt(d3 * d3 + std::norm(B1));
      energies[2 * index] = d1 - r;
      energies[2 * index + 1] = d1 + r;
      cplx v0, v1;  // null vector of H - (d1 - r)
      if (d3 >= 0.0) {
        v0 = B1;
        v1 = -(d3 + r);
      } else {
        v0 = r - d3;
        v1 = -std::conj(B1);
      }
      double nrm = std::sqrt(std::norm(v0) + std::norm(v1));
      if (nrm == 0.0) {
        v0 = 1.0;
        nrm = 1.0;
      }
      v0 /= nrm;
      v1 /= nrm;
      cplx phase = (std::abs(v0) > 1e-300) ? std::conj(v0) / std::abs(v0) : std::conj(v1) / std::abs(v1);
      v0 *= phase;
      v1 *= phase;
      u[2 * index] = std::conj(v0);
      u[2 * index + 1] = std::conj(v1);
    }
}

// ParticleOnLatticeCheckerboardLatticeSingleBandHamiltonian::EvaluateInteractionFactors
// (FTI/src/Hamiltonian/ParticleOnLatticeCheckerboardLatticeSingleBandHamiltonian.cc:126, fermions):
// nearest-neighbour A-B repulsion U (and A-A / B-B repulsion V) projected on the band u.
void CheckerboardSingleBandFactors(int nx, int ny, const std::vector<cplx>& u, double uPotential, double vPotential,
                                   bool flatBand, std::vector<int>& sumPairCount, std::vector<int>& sumPairs,
                                   std::vector<double>& factors) {
  const int ns = nx * ny;
  std::vector<std::vector<std::pair<int, int>>> pairs(ns);
  for (int kx1 = 0; kx1 < nx; ++kx1)
    for (int kx2 = 0; kx2 < nx; ++kx2)
      for (int ky1 = 0; ky1 < ny; ++ky1)
        for (int ky2 = 0; ky2 < ny; ++ky2) {
          int i1 = kx1 * ny + ky1, i2 = kx2 * ny + ky2;
          if (i1 < i2) pairs[((kx1 + kx2) % nx) * ny + ((ky1 + ky2) % ny)].push_back({i1, i2});
        }
  double factorU = 0.5 / ns;
  if (!flatBand) factorU *= uPotential;
  const double factorV = vPotential * 0.5 / ns;
  auto K = [&](int i, int c) { return c == 0 ? 2.0 * M_PI * (i / ny) / nx : 2.0 * M_PI * (i % ny) / ny; };
  auto AB = [&](int p1, int p2) {  // ComputeTwoBodyMatrixElementAB (:359)
    return 2.0 * (std::cos(0.5 * (K(p2, 0) - K(p2, 1) + K(p1, 1) - K(p1, 0))) +
                  std::cos(0.5 * (K(p2, 0) + K(p2, 1) - K(p1, 1) - K(p1, 0))));
  };
  auto AA = [&](int p1, int p2) {  // ComputeTwoBodyMatrixElementAA (:380)
    return std::polar(1.0, K(p2, 0) - K(p1, 0)) + std::polar(1.0, K(p2, 1) - K(p1, 1));
  };
  auto U = [&](int i, int alpha) { return u[2 * i + alpha]; };
  sumPairCount.assign(ns, 0);
  sumPairs.clear();
  factors.clear();
  for (int s = 0; s < ns; ++s) {
    const auto& P = pairs[s];
    sumPairCount[s] = (int)P.size();
    for (const auto& p : P) {
      sumPairs.push_back(p.first);
      sumPairs.push_back(p.second);
    }
    for (const auto& q1 : P)
      for (const auto& q2 : P) {
        const int I1 = q1.first, I2 = q1.second, I3 = q2.first, I4 = q2.second;
        // (sign, a, b, c, d | form factor momenta): conj(u_a,A) u_b,A conj(u_c,B) u_d,B
        const int t[4][7] = {{+1, I1, I3, I2, I4, I2, I4},
                             {-1, I2, I3, I1, I4, I1, I4},
                             {-1, I1, I4, I2, I3, I2, I3},
                             {+1, I2, I4, I1, I3, I1, I3}};
        cplx f = 0.0;
        for (const auto& r : t) {
          f += (double)r[0] * factorU * (std::conj(U(r[1], 0)) * U(r[2], 0) * std::conj(U(r[3], 1)) * U(r[4], 1)) *
               AB(r[5], r[6]);
          if (factorV != 0.0)
            f += (double)r[0] * factorV *
                 (std::conj(U(r[1], 0)) * U(r[2], 0) * std::conj(U(r[3], 0)) * U(r[4], 0) +
                  std::conj(U(r[1], 1)) * U(r[2], 1) * std::conj(U(r[3], 1)) * U(r[4], 1)) *
                 AA(r[5], r[6]);
        }
        f *= -2.0;
        factors.push_back(f.real());
        factors.push_back(f.imag());
      }
  }
}

// eigenvalues of the complex Hermitian block tridiagonal matrix, through its real 2n x 2n form
// (every eigenvalue appears twice there)
void BlockTridiagonalEigen(const std::vector<std::vector<cplx>>& A, const std::vector<std::vector<cplx>>& B, int k,
                           int m, std::vector<double>& evals, std::vector<cplx>* evecs) {
  if (evecs) throw std::invalid_argument("complex Ritz vectors are not implemented");
  const int n = m * k, n2 = 2 * n;
  std::vector<double> T((size_t)n2 * n2, 0.0);
  auto put = [&](int i, int j, cplx v) {  // complex entry (i, j) -> real 2 x 2 block
    T[(size_t)(2 * i) * n2 + 2 * j] = v.real();
    T[(size_t)(2 * i) * n2 + 2 * j + 1] = -v.imag();
    T[(size_t)(2 * i + 1) * n2 + 2 * j] = v.imag();
    T[(size_t)(2 * i + 1) * n2 + 2 * j + 1] = v.real();
  };
  for (int j = 0; j < m; ++j)
    for (int p = 0; p < k; ++p)
      for (int q = 0; q < k; ++q) {
        put(j * k + p, j * k + q, A[j][(size_t)p * k + q]);
        if (j + 1 < m) {
          cplx b = B[j][(size_t)p * k + q];  // T[j+1][j] block
          put((j + 1) * k + p, j * k + q, b);
          put(j * k + q, (j + 1) * k + p, std::conj(b));
        }
      }
  std::vector<double> d, e, w;
  BandToTridiagonal(n2, 2 * k + 1, T, d, e, nullptr);
  TridiagonalEigen(n2, d, e, w, nullptr);
  evals.resize(n);
  for (int i = 0; i < n; ++i) evals[i] = w[2 * i];
}

LatticeSectorsBlockLanczos::LatticeSectorsBlockLanczos(GPUArchitecture* architecture, int nbrFermions, int nx, int ny,
                                                       const std::vector<int>& kx, const std::vector<int>& ky,
                                                       const std::vector<int>& sumPairCount,
                                                       const std::vector<int>& sumPairs,
                                                       const std::vector<double>& factors, int nbrEigenvalue,
                                                       int blockSize, int maxIter, double precision,
                                                       size_t memoryBytes)
    : Arch(architecture),
      NbrFermions(nbrFermions),
      NbrSiteX(nx),
      NbrSiteY(ny),
      NbrEigenvalue(nbrEigenvalue),
      BlockSize(blockSize),
      MaxIter(maxIter),
      Precision(precision),
      MemoryBytes(memoryBytes ? memoryBytes : SIZE_MAX),
      Kx(kx),
      Ky(ky),
      SumPairCount(sumPairCount),
      SumPairs(sumPairs),
      Factors(factors) {
  if (BlockSize != 1 && BlockSize != 2 && BlockSize != 4 && BlockSize != 8)
    throw std::invalid_argument("complex block size must be 1, 2, 4 or 8");
}

std::vector<SectorResult> LatticeSectorsBlockLanczos::Solve(uint64_t seed, double* seconds) {
  auto t0 = std::chrono::steady_clock::now();
  const int ns = (int)Kx.size();
  std::vector<SectorResult> out(ns);
  const int k = BlockSize;
  // sector dimensions: counted on the host for momentum sectors; the orbits of the torus with
  // translations come from a set of all sectors
  std::vector<int64_t> dims(ns);
  if (!TorusTranslations) {
    for (int s = 0; s < ns; ++s) dims[s] = dh_momentum_sector_dimension(NbrFermions, NbrSiteX, NbrSiteY, Kx[s], Ky[s]);
  } else {
    dh_sector_set* S = CreateSet(Kx, Ky);
    std::vector<int64_t> off(ns + 1);
    dh_sector_set_offsets(S, off.data());
    for (int s = 0; s < ns; ++s) dims[s] = off[s + 1] - off[s];
    dh_sector_set_destroy(S);
  }
  for (int s = 0; s < ns; ++s) {
    out[s].dimension = dims[s];
    if (dims[s] <= 0) throw std::invalid_argument("empty momentum sector");
  }
  // waves whose expected Krylov storage fits the device (RunSectorWaves)
  Waves = RunSectorWaves(
      dims, k, 2 * k, NbrEigenvalue, MaxIter, MemoryBytes, out,
      [&](const std::vector<int>& wave, int blocks) { SolveWave(wave, seed, out, blocks); },
      [&] { return DeviceFreeBytes(Arch); });
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return out;
}

dh_sector_set* LatticeSectorsBlockLanczos::CreateSet(const std::vector<int>& kx, const std::vector<int>& ky) const {
  dh_context* c = Arch->Context();
  dh_sector_set* S = nullptr;
  if (TorusTranslations)
    Arch->Check(dh_torus_translation_sector_set_create(c, &S, NbrFermions, NbrSiteY, kx.data(), ky.data(),
                                                       (int)kx.size()),
                "torus sector set");
  else
    Arch->Check(dh_momentum_sector_set_create(c, &S, NbrFermions, NbrSiteX, NbrSiteY, kx.data(), ky.data(),
                                              (int)kx.size()),
                "momentum sector set");
  return S;
}

void LatticeSectorsBlockLanczos::SolveWave(const std::vector<int>& sectors, uint64_t seed,
                                           std::vector<SectorResult>& out, int expectedBlocks) {
  dh_context* c = Arch->Context();
  const int k = BlockSize;
  const int na = (int)sectors.size();
  std::vector<int> kx, ky;
  for (int s : sectors) {
    kx.push_back(Kx[s]);
    ky.push_back(Ky[s]);
  }
  SectorSolveResources guard(Arch);
  guard.set = CreateSet(kx, ky);
  dh_sector_set* S = guard.set;
  Arch->Check(dh_sector_set_lattice_factors(S, NbrSiteX * NbrSiteY, SumPairCount.data(), SumPairs.data(),
                                            Factors.data(), 0.0),
              "lattice factors");
  std::vector<int64_t> off(na + 1);
  dh_sector_set_offsets(S, off.data());
  const int64_t total = off.back();
  // small sets: four lanes per row for enough wavefronts; large ones: the fastest shapes of N=10
  // on 6 x 5 sites, 30 sectors, measured on MI355X: k <= 2 two lanes per row and 512-thread
  // workgroups (39.9 ms at k = 1), k > 2 one lane per row and 256-thread workgroups (84.5 ms at
  // k = 4, against 90.4 ms with two lanes and 512 threads)
  if (Launch.lanesPerRow > 0)
    Arch->Check(dh_sector_set_launch(S, Launch.lanesPerRow, Launch.blockSize), "launch");
  else if (total < 4ll * 256 * 256)
    Arch->Check(dh_sector_set_launch(S, 4, 256), "launch");
  else if (k <= 2)
    Arch->Check(dh_sector_set_launch(S, 2, 512), "launch");
  else
    Arch->Check(dh_sector_set_launch(S, 1, 256), "launch");
  std::vector<int64_t> dims(na);
  for (int i = 0; i < na; ++i) dims[i] = off[i + 1] - off[i];
  std::vector<int> blockCap;
  const int maxBlocks = SectorBlockCaps(dims, k, MaxIter, blockCap);
  const int64_t blockDoubles = total * 2 * k;
  // the blocks the planner expects the wave to need, or the usual start when it expects fewer
  const int capacity = std::min(maxBlocks, std::max(InitialKrylovBlocks(k, maxBlocks), expectedBlocks));
  guard.basis.push_back(Arch->AllocateMemory((int64_t)(capacity + 2) * blockDoubles));
  SectorSetOps ops(Arch, S, k, 2 * k, total);
  SegmentedBlockLanczos<cplx> lanczos(&ops, dims, k, maxBlocks, guard.basis[0], blockDoubles);
  lanczos.SetGrowth(capacity, [&](int& blocks, int need) {
    return GrowKrylovStorage(Arch, guard.basis, blocks, need, blockDoubles);
  });
  std::vector<SectorResult*> res;
  for (int s : sectors) res.push_back(&out[s]);
  RunSectorLanczos(lanczos, blockCap, NbrEigenvalue, Precision, seed, res);
  Arch->Check(dh_synchronize(c), "sync");
}

}  // namespace dhh
