// Spinless fermions in one Chern band on a lattice, in momentum space: complex FP64 H.v.
//
// Reference path: FermionOnSquareLatticeMomentumSpace (FQHE/src/HilbertSpace/
// FermionOnSquareLatticeMomentumSpace.cc:76, GenerateStates :290) with the FermionOnSphere
// bit, sign and search conventions, and ParticleOnLatticeTimeReversalBreakingSingleBandHamiltonian::
// EvaluateMNTwoBodyAddMultiplyComponent (FQHE/src/Hamiltonian/ParticleOnLatticeTimeReversalBreakingSingleBandHamiltonian.h:370):
// for every momentum-sum sector j, every annihilated pair i1 = (a < b) of that sum and every
// created pair i2 = (c < d) of the same sum, target += sign * F[j][i1, i2] * x.
// The kernel gathers: H is Hermitian, so row r of H is the conjugate of the terms generated
// from state r, y[r] = sum_terms sign * conj(F) * x[target].
//
// Orbital o = kx * Ny + ky; a sector is a total momentum class (Kx mod Nx, Ky mod Ny).  The
// basis is unranked on the device and FindStateIndex is the combinatorial rank with the
// momentum class taking the place of the orbital-index sum of the sphere.
#include <algorithm>
#include <cstring>

#include "dh_index.h"
#include "dh_internal.h"

namespace {
struct Classes {
  int nx, ny;
  int add(int a, int b) const { return (((a / ny + b / ny) % nx) * ny) + ((a % ny + b % ny) % ny); }
  int sub(int a, int b) const { return (((a / ny - b / ny + nx) % nx) * ny) + ((a % ny - b % ny + ny) % ny); }
};

// T[(j * (N+1) + n) * C + c]: number of n-subsets of orbitals {0..j-1} with momentum class c
std::vector<int64_t> class_table(int nOrb, int N, const Classes& cl) {
  const int C = cl.nx * cl.ny;
  std::vector<int64_t> t((size_t)(nOrb + 1) * (N + 1) * C, 0);
  auto idx = [&](int j, int n, int c) { return ((size_t)j * (N + 1) + n) * C + c; };
  t[idx(0, 0, 0)] = 1;
  for (int j = 0; j < nOrb; ++j)
    for (int n = 0; n <= N; ++n)
      for (int c = 0; c < C; ++c) {
        int64_t v = t[idx(j, n, c)];
        if (n >= 1) v += t[idx(j, n - 1, cl.sub(c, j))];
        t[idx(j + 1, n, c)] = v;
      }
  return t;
}
}  // namespace

__device__ __forceinline__ int cls_sub(int a, int b, int nx, int ny) {
  return (((a / ny - b / ny + nx) % nx) * ny) + ((a % ny - b % ny + ny) % ny);
}

__global__ void lattice_unrank_kernel(uint64_t* __restrict__ states, int64_t dim, int nOrb, int N, int target,
                                      const int64_t* __restrict__ T, int nx, int ny) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dim) return;
  const int C = nx * ny;
  int64_t rem = i;
  int n = N, c = target;
  uint64_t state = 0;
  for (int j = nOrb - 1; j >= 0 && n > 0; --j) {
    int rest = cls_sub(c, j, nx, ny);
    int64_t cnt = T[((int64_t)j * (N + 1) + (n - 1)) * C + rest];
    if (rem < cnt) {
      state |= (1ull << j);
      --n;
      c = rest;
    } else {
      rem -= cnt;
    }
  }
  states[i] = state;
}

// FermionOnSquareLatticeMomentumSpace::EvaluateHilbertSpaceDimension
// (FQHE/src/HilbertSpace/FermionOnSquareLatticeMomentumSpace.cc:290): the number of N-subsets of
// the nx * ny momenta with total momentum (kx, ky), from the class table; no device work.
extern "C" int64_t dh_momentum_sector_dimension(int N, int nx, int ny, int kx, int ky) {
  const int nOrb = nx * ny;
  if (nx <= 0 || ny <= 0 || N <= 0 || N > nOrb || nOrb > 62 || kx < 0 || kx >= nx || ky < 0 || ky >= ny) return -1;
  Classes cl{nx, ny};
  const std::vector<int64_t> T = class_table(nOrb, N, cl);
  return T[((size_t)nOrb * (N + 1) + N) * nOrb + (kx * ny + ky)];
}

extern "C" int dh_momentum_sector_set_create(dh_context* c, dh_sector_set** out, int N, int nx, int ny,
                                             const int* kx, const int* ky, int nbrSectors) {
  if (!c || !out || !kx || !ky || nbrSectors <= 0) return DH_E_ARG;
  *out = nullptr;
  if (c->closed) return dh_fail(c, DH_E_STATE, "context destroyed");
  const int nOrb = nx * ny;
  if (nx <= 0 || ny <= 0 || N <= 0 || N > nOrb) return dh_fail(c, DH_E_ARG, "need 0 < nbr_fermions <= Nx * Ny");
  if (nOrb > 62) return dh_fail(c, DH_E_RANGE, "Nx * Ny must be <= 62 (64-bit states)");
  hipSetDevice(c->device);
  Classes cl{nx, ny};
  const int C = nOrb;
  std::vector<int64_t> T = class_table(nOrb, N, cl);
  auto Tat = [&](int j, int n, int cc) -> int64_t {
    if (j < 0 || j > nOrb || n < 0 || n > N) return 0;
    return T[((size_t)j * (N + 1) + n) * C + cc];
  };
  dh_sector_set* S = new dh_sector_set();
  S->ctx = c;
  S->kind = DH_SECTORS_MOMENTUM;
  S->nbrFermions = N;
  S->nbrOrbitals = nOrb;
  S->lzMax = nOrb;
  S->nbrSiteX = nx;
  S->nbrSiteY = ny;
  S->nbrSectors = nbrSectors;
  S->offsets.assign(nbrSectors + 1, 0);
  std::vector<int> target(nbrSectors);
  for (int i = 0; i < nbrSectors; ++i) {
    if (kx[i] < 0 || kx[i] >= nx || ky[i] < 0 || ky[i] >= ny) {
      delete S;
      return dh_fail(c, DH_E_ARG, "momentum sector outside the Brillouin zone");
    }
    target[i] = kx[i] * ny + ky[i];
    int64_t d = Tat(nOrb, N, target[i]);
    if (d <= 0 || d >= (1ll << 31)) {
      delete S;
      return dh_fail(c, DH_E_ARG, "a sector is empty or larger than 2^31");
    }
    S->kx.push_back(kx[i]);
    S->ky.push_back(ky[i]);
    S->offsets[i + 1] = S->offsets[i] + d;
  }
  S->totalDim = S->offsets[nbrSectors];
  S->rankBits = dh_rank_split(nOrb);
  S->loSize = 1ll << S->rankBits;
  S->hiSize = 1ll << (nOrb - S->rankBits);
  std::vector<uint32_t> lo(S->loSize, 0), hi((size_t)S->hiSize * nbrSectors, 0);
  for (uint64_t l = 0; l < (uint64_t)S->loSize; ++l) {
    int n = 0, cc = 0;
    int64_t acc = 0;
    for (uint64_t r = l; r; r &= r - 1) {
      int j = __builtin_ctzll(r);
      ++n;
      cc = cl.add(cc, j);
      acc += Tat(j, n, cc);
    }
    lo[l] = (uint32_t)acc;
  }
  dh_parallel_for(nbrSectors, [&](int sec) {
    for (uint64_t h = 0; h < (uint64_t)S->hiSize; ++h) {
      int popH = __builtin_popcountll(h), clsH = 0;
      for (uint64_t r = h; r; r &= r - 1) clsH = cl.add(clsH, S->rankBits + __builtin_ctzll(r));
      int n = N - popH;
      if (n < 0) continue;
      int cc = cl.sub(target[sec], clsH);
      int64_t acc = 0;
      for (uint64_t r = h; r; r &= r - 1) {
        int j = S->rankBits + __builtin_ctzll(r);
        ++n;
        cc = cl.add(cc, j);
        acc += Tat(j, n, cc);
      }
      hi[(size_t)sec * S->hiSize + h] = (uint32_t)acc;
    }
  });
  const std::vector<uint16_t> lo16 = dh_rank_lo16(lo);
  int rc;
  int64_t* d_T = nullptr;
  if ((rc = dh_alloc(c, (void**)&d_T, sizeof(int64_t) * T.size())) ||
      (rc = dh_alloc(c, (void**)&S->d_states, sizeof(uint64_t) * S->totalDim)) ||
      (rc = dh_alloc(c, (void**)&S->d_rankLo, sizeof(uint32_t) * lo.size())) ||
      (rc = dh_alloc(c, (void**)&S->d_rankHi, sizeof(uint32_t) * hi.size())) ||
      (!lo16.empty() && (rc = dh_alloc(c, (void**)&S->d_rankLo16, sizeof(uint16_t) * lo16.size()))) ||
      (rc = dh_alloc(c, (void**)&S->d_offsets, sizeof(int64_t) * (nbrSectors + 1)))) {
    if (d_T) hipFree(d_T);
    dh_sector_set_destroy(S);
    return rc;
  }
  hipMemcpyAsync(d_T, T.data(), sizeof(int64_t) * T.size(), hipMemcpyHostToDevice, c->stream);
  hipMemcpyAsync(S->d_rankLo, lo.data(), sizeof(uint32_t) * lo.size(), hipMemcpyHostToDevice, c->stream);
  hipMemcpyAsync(S->d_rankHi, hi.data(), sizeof(uint32_t) * hi.size(), hipMemcpyHostToDevice, c->stream);
  if (S->d_rankLo16)
    hipMemcpyAsync(S->d_rankLo16, lo16.data(), sizeof(uint16_t) * lo16.size(), hipMemcpyHostToDevice, c->stream);
  hipMemcpyAsync(S->d_offsets, S->offsets.data(), sizeof(int64_t) * (nbrSectors + 1), hipMemcpyHostToDevice,
                 c->stream);
  for (int sec = 0; sec < nbrSectors; ++sec) {
    int64_t d = S->offsets[sec + 1] - S->offsets[sec];
    hipLaunchKernelGGL(lattice_unrank_kernel, dim3((unsigned)((d + 255) / 256)), dim3(256), 0, c->stream,
                       S->d_states + S->offsets[sec], d, nOrb, N, target[sec], d_T, nx, ny);
  }
  hipError_t e = hipStreamSynchronize(c->stream);
  hipFree(d_T);
  if (e != hipSuccess) {
    dh_sector_set_destroy(S);
    return dh_fail(c, DH_E_HIP, std::string("momentum sector set: ") + hipGetErrorString(e));
  }
  *out = S;
  return DH_OK;
}

// Free-pair tables of a momentum set.  For momentum sum j = (jx, jy) the partner of orbital
// m = kx Ny + ky is P_j(m) = ((jx - kx) mod Nx) Ny + (jy - ky) mod Ny.  The mirror of the free mask f
// (bit m set when P_j(m) is free) is f reversed within the N orbitals, which maps (kx, ky) to
// (Nx-1-kx, Ny-1-ky), then rotated by ax = (Nx-1-jx) mod Nx blocks of Ny bits and by
// ay = (Ny-1-jy) mod Ny bits inside every block.  Per sum: the rotation in bits ax Ny and ay, the
// masks of the positions whose ky + ay stays below Ny (taken from g >> ay) or wraps (from
// g << (Ny - ay)), the mask of the orbitals that are the smaller of their pair, and per orbital the
// pair's position in the sum's list with its partner.
static int lattice_free_tables(dh_sector_set* S, const std::vector<int>& info) {
  dh_context* c = S->ctx;
  const int nx = S->nbrSiteX, ny = S->nbrSiteY, N = S->nbrOrbitals, J = nx * ny;
  std::vector<uint64_t> masks(3 * (size_t)J, 0);
  std::vector<int> shifts(2 * (size_t)J, 0), pp((size_t)J * N, -1);
  for (int j = 0; j < J; ++j) {
    const int jx = j / ny, jy = j % ny;
    const int ax = (nx - 1 - jx) % nx, ay = (ny - 1 - jy) % ny;
    shifts[2 * j] = ax * ny;
    shifts[2 * j + 1] = ay;
    for (int m = 0; m < N; ++m) {
      const int kx = m / ny, ky = m % ny;
      if (ky < ny - ay)
        masks[3 * j] |= 1ull << m;
      else
        masks[3 * j + 1] |= 1ull << m;
      const int partner = ((jx - kx + nx) % nx) * ny + ((jy - ky + ny) % ny);
      if (m < partner) {
        masks[3 * j + 2] |= 1ull << m;
        const int id = partner * (partner - 1) / 2 + m;
        if ((info[id] >> 16) != j) return dh_fail(c, DH_E_STATE, "momentum pair table inconsistent");
        pp[(size_t)j * N + m] = ((info[id] & 0xffff) << 8) | partner;
      }
    }
  }
  for (void* p : {(void*)S->d_freeMask, (void*)S->d_freeShift, (void*)S->d_partnerPos})
    if (p) hipFree(p);
  S->d_freeMask = nullptr;
  S->d_freeShift = nullptr;
  S->d_partnerPos = nullptr;
  int rc;
  if ((rc = dh_alloc(c, (void**)&S->d_freeMask, sizeof(uint64_t) * masks.size())) ||
      (rc = dh_alloc(c, (void**)&S->d_freeShift, sizeof(int) * shifts.size())) ||
      (rc = dh_alloc(c, (void**)&S->d_partnerPos, sizeof(int) * pp.size())))
    return rc;
  DH_HIP(c, hipMemcpyAsync(S->d_freeMask, masks.data(), sizeof(uint64_t) * masks.size(), hipMemcpyHostToDevice,
                           c->stream));
  DH_HIP(c, hipMemcpyAsync(S->d_freeShift, shifts.data(), sizeof(int) * shifts.size(), hipMemcpyHostToDevice,
                           c->stream));
  DH_HIP(c, hipMemcpyAsync(S->d_partnerPos, pp.data(), sizeof(int) * pp.size(), hipMemcpyHostToDevice, c->stream));
  DH_HIP(c, hipStreamSynchronize(c->stream));
  return DH_OK;
}

extern "C" int dh_sector_set_lattice_factors(dh_sector_set* S, int nbrSums, const int* sumPairCount,
                                             const int* sumPairs, const double* factors, double shift) {
  if (!S || !sumPairCount || !sumPairs || !factors) return DH_E_ARG;
  dh_context* c = S->ctx;
  if (S->kind != DH_SECTORS_MOMENTUM && S->kind != DH_SECTORS_TORUS_TRANSLATIONS)
    return dh_fail(c, DH_E_ARG, "not a momentum or torus sector set");
  const int nOrb = S->nbrOrbitals;
  if (nbrSums != S->nbrSiteX * S->nbrSiteY) return dh_fail(c, DH_E_ARG, "nbr_sums must be Nx * Ny");
  const int nPairs = nOrb * (nOrb - 1) / 2;
  std::vector<int> info(nPairs, -1), pairOff(nbrSums + 1, 0), packed;
  std::vector<int64_t> facOff(nbrSums + 1, 0);
  int pos = 0;
  for (int j = 0; j < nbrSums; ++j) {
    int n = sumPairCount[j];
    if (n < 0 || n > 65535) return dh_fail(c, DH_E_ARG, "pair count");
    pairOff[j + 1] = pairOff[j] + n;
    facOff[j + 1] = facOff[j] + (int64_t)n * n;
    for (int i = 0; i < n; ++i, ++pos) {
      int a = sumPairs[2 * pos], b = sumPairs[2 * pos + 1];
      if (a < 0 || b <= a || b >= nOrb) return dh_fail(c, DH_E_ARG, "pairs must be (a < b) orbitals");
      int cls = (((a / S->nbrSiteY + b / S->nbrSiteY) % S->nbrSiteX) * S->nbrSiteY) +
                ((a % S->nbrSiteY + b % S->nbrSiteY) % S->nbrSiteY);
      if (cls != j) return dh_fail(c, DH_E_ARG, "pair listed under the wrong momentum sum");
      int id = b * (b - 1) / 2 + a;
      if (info[id] != -1) return dh_fail(c, DH_E_ARG, "pair listed twice");
      info[id] = (j << 16) | i;
      packed.push_back(a | (b << 8));
    }
  }
  for (int id = 0; id < nPairs; ++id)
    if (info[id] == -1) return dh_fail(c, DH_E_ARG, "every pair a < b must be listed once");
  for (void* p : {(void*)S->d_pairInfo, (void*)S->d_sumPairOff, (void*)S->d_sumPairs, (void*)S->d_sumFactorOff,
                  (void*)S->d_cfactors})
    if (p) hipFree(p);
  S->d_pairInfo = S->d_sumPairOff = S->d_sumPairs = nullptr;
  S->d_sumFactorOff = nullptr;
  S->d_cfactors = nullptr;
  int rc;
  if ((rc = dh_alloc(c, (void**)&S->d_pairInfo, sizeof(int) * nPairs)) ||
      (rc = dh_alloc(c, (void**)&S->d_sumPairOff, sizeof(int) * (nbrSums + 1))) ||
      (rc = dh_alloc(c, (void**)&S->d_sumPairs, sizeof(int) * packed.size())) ||
      (rc = dh_alloc(c, (void**)&S->d_sumFactorOff, sizeof(int64_t) * (nbrSums + 1))) ||
      (rc = dh_alloc(c, (void**)&S->d_cfactors, sizeof(double) * 2 * facOff.back())))
    return rc;
  DH_HIP(c, hipMemcpyAsync(S->d_pairInfo, info.data(), sizeof(int) * nPairs, hipMemcpyHostToDevice, c->stream));
  DH_HIP(c, hipMemcpyAsync(S->d_sumPairOff, pairOff.data(), sizeof(int) * (nbrSums + 1), hipMemcpyHostToDevice,
                           c->stream));
  DH_HIP(c, hipMemcpyAsync(S->d_sumPairs, packed.data(), sizeof(int) * packed.size(), hipMemcpyHostToDevice,
                           c->stream));
  DH_HIP(c, hipMemcpyAsync(S->d_sumFactorOff, facOff.data(), sizeof(int64_t) * (nbrSums + 1), hipMemcpyHostToDevice,
                           c->stream));
  DH_HIP(c, hipMemcpyAsync(S->d_cfactors, factors, sizeof(double) * 2 * facOff.back(), hipMemcpyHostToDevice,
                           c->stream));
  DH_HIP(c, hipStreamSynchronize(c->stream));
  S->nbrSums = nbrSums;
  S->nbrPairs = nPairs;
  S->nbrComplexFactors = facOff.back();
  S->shift = shift;
  if (S->kind == DH_SECTORS_MOMENTUM) return lattice_free_tables(S, info);
  return DH_OK;
}

struct LatticeArgs {
  const uint64_t* states;
  const int64_t* offsets;
  int nbrOrbitals, nbrPairs, nbrSums, totalPairs;
  int nbrSiteY;
  const uint64_t* freeMask;
  const int* freeShift;
  const int* partnerPos;
  const int* pairInfo;
  const int* sumPairOff;
  const int* sumPairs;
  const int64_t* sumFactorOff;
  const double* factors;
  const uint32_t* rankLo;
  const uint32_t* rankHi;
  const uint16_t* rankLo16;  // the low table in 16 bits, staged in LDS (LO_LDS)
  int rankBits;
  int64_t loSize, hiSize;
  double shift;
  const int64_t* work;
  int nActive;
  int64_t totalTiles;
};

__device__ __forceinline__ int lat_tile_sector(const int64_t* sPrefix, int nActive, int64_t tile) {
  int lo = 0, hi = nActive;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (sPrefix[mid] <= tile)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// free created pairs of momentum sum j in t: bit m1 set when m1 and its partner are free and m1 is
// the smaller of the two (tables of lattice_free_tables, in LDS)
__device__ __forceinline__ uint64_t lattice_free_pairs(uint64_t t, int N, int ny, const uint64_t* __restrict__ mask,
                                                       const int* __restrict__ shift, int j) {
  const uint64_t full = N >= 64 ? ~0ull : ((1ull << N) - 1ull);
  const uint64_t f = ~t & full;
  const uint64_t R = __builtin_bitreverse64(f) >> (64 - N);
  const int r = shift[2 * j], ay = shift[2 * j + 1];
  const uint64_t g1 = r ? (((R >> r) | (R << (N - r))) & full) : R;
  const uint64_t g = ((g1 >> ay) & mask[3 * j]) | ((g1 << (ny - ay)) & mask[3 * j + 1]);
  return f & g & mask[3 * j + 2];
}

// K complex vectors per row: x and y hold 2K doubles per row, (re, im) interleaved.  FREE: the
// created pairs of each annihilated pair come from lattice_free_pairs (no iteration over a created
// pair with an occupied orbital); otherwise every listed pair of the momentum sum is tried.
template <int K, int LPR, bool LO_LDS, bool FREE>
__global__ void __launch_bounds__(1024) lattice_hv_kernel(LatticeArgs a, const double* __restrict__ x,
                                                          double* __restrict__ y, int add) {
  extern __shared__ int64_t smem64[];
  int64_t* sWork = smem64;                                    // 2 nActive + 1
  int64_t* sFacOff = sWork + 2 * a.nActive + 1;               // nbrSums + 1
  uint64_t* sMask = reinterpret_cast<uint64_t*>(sFacOff + a.nbrSums + 1);  // FREE: 3 nbrSums
  int* sInfo = reinterpret_cast<int*>(sMask + (FREE ? 3 * a.nbrSums : 0));  // nbrPairs
  int* sPairOff = sInfo + a.nbrPairs;                         // nbrSums + 1
  int* sPairs = sPairOff + a.nbrSums + 1;                     // totalPairs
  int* sShift = sPairs + a.totalPairs;                        // FREE: 2 nbrSums
  int* sPP = sShift + (FREE ? 2 * a.nbrSums : 0);             // FREE: nbrSums * nbrOrbitals
  // the low table: its 16-bit copy in LDS (LO_LDS), or the 32-bit table in global memory
  using LoT = typename std::conditional<LO_LDS, uint16_t, uint32_t>::type;
  LoT* sLo = reinterpret_cast<LoT*>(sPP + (FREE ? a.nbrSums * a.nbrOrbitals : 0));
  for (int i = threadIdx.x; i < 2 * a.nActive + 1; i += blockDim.x) sWork[i] = a.work[i];
  for (int i = threadIdx.x; i <= a.nbrSums; i += blockDim.x) {
    sFacOff[i] = a.sumFactorOff[i];
    sPairOff[i] = a.sumPairOff[i];
  }
  for (int i = threadIdx.x; i < a.nbrPairs; i += blockDim.x) sInfo[i] = a.pairInfo[i];
  for (int i = threadIdx.x; i < a.totalPairs; i += blockDim.x) sPairs[i] = a.sumPairs[i];
  if constexpr (FREE) {
    for (int i = threadIdx.x; i < 3 * a.nbrSums; i += blockDim.x) sMask[i] = a.freeMask[i];
    for (int i = threadIdx.x; i < 2 * a.nbrSums; i += blockDim.x) sShift[i] = a.freeShift[i];
    for (int i = threadIdx.x; i < a.nbrSums * a.nbrOrbitals; i += blockDim.x) sPP[i] = a.partnerPos[i];
  }
  if constexpr (LO_LDS)
    for (int i = threadIdx.x; i < a.loSize; i += blockDim.x) sLo[i] = a.rankLo16[i];
  __syncthreads();
  const LoT* lo;
  if constexpr (LO_LDS)
    lo = sLo;
  else
    lo = a.rankLo;
  const int64_t* sPrefix = sWork + a.nActive;
  const int lane = threadIdx.x % LPR;
  const int64_t rowsPerTile = blockDim.x / LPR;
  for (int64_t tile = dh_xcd_block(blockIdx.x, gridDim.x); tile < a.totalTiles; tile += gridDim.x) {
    const int j = lat_tile_sector(sPrefix, a.nActive, tile);
    const int sec = (int)sWork[j];
    const int64_t off = a.offsets[sec];
    const int64_t dim = a.offsets[sec + 1] - off;
    const uint32_t* hi = a.rankHi + (int64_t)sec * a.hiSize;
    const int64_t local = (tile - sPrefix[j]) * rowsPerTile + threadIdx.x / LPR;
    const bool active = local < dim;
    const int64_t g = off + (active ? local : 0);
    const uint64_t s = active ? a.states[g] : 0ull;
    double acc[2 * K];
#pragma unroll
    for (int v = 0; v < 2 * K; ++v) acc[v] = 0.0;
    const double* xs = x + off * 2 * K;
    if constexpr (FREE) {
      // every annihilated pair n1 < n2, then its free created pairs from the mask of lattice_free_pairs
      // (lane `lane` of LPR takes every LPR-th of them); for K <= 2 a term's multiply-add waits for
      // the next term's index loads, as in the list loop below, with the terms in the same order
      constexpr bool PIPE = K <= 2;
      double pre = 0.0, pim = 0.0;
      double2 px[K];
#pragma unroll
      for (int v = 0; v < K; ++v) px[v] = make_double2(0.0, 0.0);
      for (uint64_t r1 = s; r1; r1 &= r1 - 1) {
        const int n1 = __builtin_ctzll(r1);  // lower annihilated orbital
        for (uint64_t r2 = s & ~((2ull << n1) - 1ull); r2; r2 &= r2 - 1) {
          const int n2 = __builtin_ctzll(r2);  // higher annihilated orbital
          // AA(n1, n2): a_n2 first, then a_n1
          const int sgnA = __popcll(s >> n2) + __popcll((s ^ (1ull << n2)) >> n1);
          const uint64_t t = s ^ (1ull << n1) ^ (1ull << n2);
          const int info = sInfo[(n2 * (n2 - 1)) / 2 + n1];
          const int sum = info >> 16;
          const int np = sPairOff[sum + 1] - sPairOff[sum];
          const double* F = a.factors + 2 * (sFacOff[sum] + (int64_t)(info & 0xffff) * np);
          const int* pp = sPP + sum * a.nbrOrbitals;
          uint64_t pm = lattice_free_pairs(t, a.nbrOrbitals, a.nbrSiteY, sMask, sShift, sum);
          if constexpr (LPR > 1)
            for (int q = 0; q < lane; ++q) pm &= pm - 1;
          while (pm) {
            const int m1 = __builtin_ctzll(pm);
#pragma unroll
            for (int q = 0; q < LPR; ++q) pm &= pm - 1;  // this term, then the other lanes' next ones
            const int pk = pp[m1];
            const int m2 = pk & 0xff, i2 = pk >> 8;  // m1 < m2
            const uint64_t c1 = 1ull << m1, c2 = 1ull << m2;
            // AdAd(m1, m2), a+_m2 first, then a+_m1: popc(t >> m2) + popc((t | c2) >> m1) has the
            // parity of 1 + the occupied orbitals of t in [m1, m2)
            const int par = sgnA + 1 + __popcll(t & (c2 - c1));
            const int64_t idx = dh_find_rank(t | c1 | c2, lo, hi, a.rankBits, dim);
            const double2 f = *reinterpret_cast<const double2*>(F + 2 * i2);
            const double cr = (par & 1) ? -f.x : f.x, ci = (par & 1) ? f.y : -f.y;  // conj(F), signed
            const double* xr = xs + idx * 2 * K;
            if constexpr (PIPE) {
#pragma unroll
              for (int v = 0; v < K; ++v) {
                acc[2 * v] += pre * px[v].x - pim * px[v].y;
                acc[2 * v + 1] += pre * px[v].y + pim * px[v].x;
              }
#pragma unroll
              for (int v = 0; v < K; ++v) px[v] = *reinterpret_cast<const double2*>(xr + 2 * v);
              pre = cr;
              pim = ci;
            } else {
#pragma unroll
              for (int v = 0; v < K; ++v) {
                double2 p = *reinterpret_cast<const double2*>(xr + 2 * v);
                acc[2 * v] += cr * p.x - ci * p.y;
                acc[2 * v + 1] += cr * p.y + ci * p.x;
              }
            }
          }
        }
      }
      if constexpr (PIPE) {
#pragma unroll
        for (int v = 0; v < K; ++v) {
          acc[2 * v] += pre * px[v].x - pim * px[v].y;
          acc[2 * v + 1] += pre * px[v].y + pim * px[v].x;
        }
      }
    } else {
      // K <= 2: a term's multiply-add waits until the next term's index loads are issued, so the
      // gather of one term and the index loads of the next are in flight together (terms are still
      // added in order)
      constexpr bool PIPE = K <= 2;
      double pre = 0.0, pim = 0.0;
      double2 px[K];
#pragma unroll
      for (int v = 0; v < K; ++v) px[v] = make_double2(0.0, 0.0);
      for (uint64_t r1 = s; r1; r1 &= r1 - 1) {
        const int n1 = __builtin_ctzll(r1);  // lower annihilated orbital
        for (uint64_t r2 = s & ~((2ull << n1) - 1ull); r2; r2 &= r2 - 1) {
          const int n2 = __builtin_ctzll(r2);  // higher annihilated orbital
          // AA(n1, n2): a_n2 first, then a_n1
          const int sgnA = __popcll(s >> n2) + __popcll((s ^ (1ull << n2)) >> n1);
          const uint64_t t = s ^ (1ull << n1) ^ (1ull << n2);
          const int info = sInfo[(n2 * (n2 - 1)) / 2 + n1];
          const int sum = info >> 16, i1 = info & 0xffff;
          const int p0 = sPairOff[sum], np = sPairOff[sum + 1] - p0;
          const double* F = a.factors + 2 * (sFacOff[sum] + (int64_t)i1 * np);
          for (int i2 = lane; i2 < np; i2 += LPR) {
            const int pk = sPairs[p0 + i2];
            const int m1 = pk & 0xff, m2 = pk >> 8;  // m1 < m2
            const uint64_t cm = (1ull << m1) | (1ull << m2);
            if (t & cm) continue;
            // AdAd(m1, m2): a+_m2 first, then a+_m1
            const int sg = sgnA + __popcll(t >> m2) + __popcll((t | (1ull << m2)) >> m1);
            const int64_t idx = dh_find_rank(t | cm, lo, hi, a.rankBits, dim);
            double cr = F[2 * i2], ci = -F[2 * i2 + 1];  // conj(F)
            if (sg & 1) {
              cr = -cr;
              ci = -ci;
            }
            const double* xr = xs + idx * 2 * K;
            if constexpr (PIPE) {
#pragma unroll
              for (int v = 0; v < K; ++v) {
                acc[2 * v] += pre * px[v].x - pim * px[v].y;
                acc[2 * v + 1] += pre * px[v].y + pim * px[v].x;
              }
#pragma unroll
              for (int v = 0; v < K; ++v) px[v] = *reinterpret_cast<const double2*>(xr + 2 * v);
              pre = cr;
              pim = ci;
            } else {
#pragma unroll
              for (int v = 0; v < K; ++v) {
                double2 p = *reinterpret_cast<const double2*>(xr + 2 * v);
                acc[2 * v] += cr * p.x - ci * p.y;
                acc[2 * v + 1] += cr * p.y + ci * p.x;
              }
            }
          }
        }
      }
      if constexpr (PIPE) {
#pragma unroll
        for (int v = 0; v < K; ++v) {
          acc[2 * v] += pre * px[v].x - pim * px[v].y;
          acc[2 * v + 1] += pre * px[v].y + pim * px[v].x;
        }
      }
    }
    if constexpr (LPR > 1) {
#pragma unroll
      for (int v = 0; v < 2 * K; ++v)
        for (int o = LPR / 2; o > 0; o >>= 1) acc[v] += __shfl_xor(acc[v], o, LPR);
    }
    if (active && lane == 0) {
      const double* xr = x + g * 2 * K;
      double* yr = y + g * 2 * K;
#pragma unroll
      for (int v = 0; v < 2 * K; ++v) yr[v] = (add ? yr[v] : 0.0) + acc[v] + a.shift * xr[v];
    }
  }
}

template <int K, int LPR>
static int lattice_launch(dh_sector_set* S, const double* x, double* y, int add, const unsigned char* active) {
  dh_context* c = S->ctx;
  int64_t rowsPerTile = S->blockSize / LPR;
  int na;
  int64_t tiles;
  int rc = dh_sectors_upload_work(S, active, rowsPerTile, &na, &tiles);
  if (rc) return rc;
  if (tiles == 0) return DH_OK;
  LatticeArgs a;
  a.states = S->d_states;
  a.offsets = S->d_offsets;
  a.nbrOrbitals = S->nbrOrbitals;
  a.nbrPairs = S->nbrPairs;
  a.nbrSums = S->nbrSums;
  a.totalPairs = S->nbrPairs;
  a.pairInfo = S->d_pairInfo;
  a.sumPairOff = S->d_sumPairOff;
  a.sumPairs = S->d_sumPairs;
  a.sumFactorOff = S->d_sumFactorOff;
  a.factors = S->d_cfactors;
  a.rankLo = S->d_rankLo;
  a.rankLo16 = S->d_rankLo16;
  a.rankHi = S->d_rankHi;
  a.rankBits = S->rankBits;
  a.loSize = S->loSize;
  a.hiSize = S->hiSize;
  a.shift = S->shift;
  a.work = S->d_work;
  a.nActive = na;
  a.totalTiles = tiles;
  a.nbrSiteY = S->nbrSiteY;
  a.freeMask = S->d_freeMask;
  a.freeShift = S->d_freeShift;
  a.partnerPos = S->d_partnerPos;
  // DH_INDEX_RANK: the free-pair masks for one vector (35.7 against 40.3 ms for the 30 sectors of
  // N=10 on 6 x 5 sites on MI355X), the pair list above (equal at k = 2, faster at k = 4)
  const bool free = (S->rowMode == DH_INDEX_RANK_FREE || (S->rowMode == DH_INDEX_RANK && K == 1)) && S->d_freeMask;
  size_t base = 8 * (2 * (size_t)na + 1) + 8 * (size_t)(S->nbrSums + 1) + 4 * (size_t)S->nbrPairs +
                4 * (size_t)(S->nbrSums + 1) + 4 * (size_t)S->nbrPairs;
  if (free) base += 8 * 3 * (size_t)S->nbrSums + 4 * 2 * (size_t)S->nbrSums + 4 * (size_t)S->nbrSums * S->nbrOrbitals;
  base = (base + 7) & ~size_t(7);
  size_t withLo = base + 2 * (size_t)S->loSize;
  bool loLds = S->d_rankLo16 && withLo <= 64 * 1024;
  size_t lds = loLds ? withLo : base;
  if (lds > 160 * 1024) return dh_fail(c, DH_E_RANGE, "staged tables exceed LDS");
  auto launch = [&](auto kern) -> int {
    if (lds > 64 * 1024)
      DH_HIP(c, hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const int perCU = dh_blocks_per_cu(c, kern, S->blockSize, lds);
    int64_t grid = std::max<int64_t>(1, std::min<int64_t>(tiles, (int64_t)c->nbrCU * perCU));
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(S->blockSize), lds, c->stream, a, x, y, add);
    DH_HIP(c, hipGetLastError());
    return DH_OK;
  };
  if (free) return loLds ? launch(lattice_hv_kernel<K, LPR, true, true>) : launch(lattice_hv_kernel<K, LPR, false, true>);
  if (loLds) return launch(lattice_hv_kernel<K, LPR, true, false>);
  return launch(lattice_hv_kernel<K, LPR, false, false>);
}

template <int K>
static int lattice_lpr(dh_sector_set* S, const double* x, double* y, int add, const unsigned char* active) {
  switch (S->lanesPerRow) {
    case 1: return lattice_launch<K, 1>(S, x, y, add, active);
    case 2: return lattice_launch<K, 2>(S, x, y, add, active);
    case 4: return lattice_launch<K, 4>(S, x, y, add, active);
    default: return lattice_launch<K, 8>(S, x, y, add, active);
  }
}

// Matrix elements per row of a momentum sector (the terms of lattice_hv_kernel whose target is in
// the space), summed per sector: one row per thread, a workgroup sum, one atomic add per workgroup.
__global__ void __launch_bounds__(256) lattice_elements_kernel(const uint64_t* __restrict__ states, int64_t dim,
                                                               const int* __restrict__ pairInfo,
                                                               const int* __restrict__ sumPairOff,
                                                               const int* __restrict__ sumPairs,
                                                               unsigned long long* __restrict__ total) {
  __shared__ unsigned long long red[256 / 64];
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long cnt = 0;
  if (row < dim) {
    const uint64_t s = states[row];
    for (uint64_t r1 = s; r1; r1 &= r1 - 1) {
      const int n1 = __builtin_ctzll(r1);
      for (uint64_t r2 = s & ~((2ull << n1) - 1ull); r2; r2 &= r2 - 1) {
        const int n2 = __builtin_ctzll(r2);
        const uint64_t t = s ^ (1ull << n1) ^ (1ull << n2);
        const int sum = pairInfo[(n2 * (n2 - 1)) / 2 + n1] >> 16;
        for (int p = sumPairOff[sum]; p < sumPairOff[sum + 1]; ++p) {
          const int pk = sumPairs[p];
          cnt += (t & ((1ull << (pk & 0xff)) | (1ull << (pk >> 8)))) ? 0ull : 1ull;
        }
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    atomicAdd(total, t);
  }
}

int dh_lattice_count_elements(dh_sector_set* S, int64_t* counts) {
  dh_context* c = S->ctx;
  if (!S->d_cfactors) return dh_fail(c, DH_E_STATE, "lattice factors not set");
  unsigned long long* d_total = nullptr;
  int rc = dh_alloc(c, (void**)&d_total, sizeof(unsigned long long) * S->nbrSectors);
  if (rc) return rc;
  std::vector<unsigned long long> zero(S->nbrSectors, 0ull);
  hipMemcpyAsync(d_total, zero.data(), sizeof(unsigned long long) * S->nbrSectors, hipMemcpyHostToDevice, c->stream);
  for (int sec = 0; sec < S->nbrSectors; ++sec) {
    const int64_t off = S->offsets[sec], dim = S->offsets[sec + 1] - off;
    if (dim <= 0) continue;
    hipLaunchKernelGGL(lattice_elements_kernel, dim3((unsigned)((dim + 255) / 256)), dim3(256), 0, c->stream,
                       S->d_states + off, dim, S->d_pairInfo, S->d_sumPairOff, S->d_sumPairs, d_total + sec);
  }
  std::vector<unsigned long long> total(S->nbrSectors);
  hipMemcpyAsync(total.data(), d_total, sizeof(unsigned long long) * S->nbrSectors, hipMemcpyDeviceToHost, c->stream);
  hipError_t e = hipStreamSynchronize(c->stream);
  hipFree(d_total);
  if (e != hipSuccess) return dh_fail(c, DH_E_HIP, hipGetErrorString(e));
  for (int sec = 0; sec < S->nbrSectors; ++sec) counts[sec] = (int64_t)total[sec];
  return DH_OK;
}

int dh_lattice_sectors_hv(dh_sector_set* S, const double* x, double* y, int k, int add, const unsigned char* active) {
  if (!S->d_cfactors) return dh_fail(S->ctx, DH_E_STATE, "lattice factors not set");
  switch (k) {
    case 1: return lattice_lpr<1>(S, x, y, add, active);
    case 2: return lattice_lpr<2>(S, x, y, add, active);
    case 4: return lattice_lpr<4>(S, x, y, add, active);
    case 8: return lattice_lpr<8>(S, x, y, add, active);
    default: return dh_fail(S->ctx, DH_E_ARG, "k must be 1, 2, 4 or 8 for complex vectors");
  }
}
