// Many Lz sectors of one fermions-on-the-sphere problem in single launches.
//
// The reference solves the sectors one after the other (FQHESphereFermionsTwoBodyGeneric.cc:270
// loops over Lz, each with its own FermionOnSphere, Hamiltonian and Lanczos).  Here the
// sectors of uneven dimension share one concatenated basis, one factor table and one low-bit
// rank table; a product over every active sector is one persistent launch whose workgroups
// stride over (sector, row-tile) work items, and the Lanczos vector operations are segmented
// by sector so that every sector keeps its own Krylov space.
//
// Block vectors of the set: per sector a [D_s][k] block, sectors concatenated in order, so
// global row g of sector s is offset_s + local row.
#include <algorithm>
#include <cstring>

#include "dh_index.h"
#include "dh_internal.h"
#include "dh_sphere_row.h"

// device-side subset-count table (same recursion as the single-sector basis)
static std::vector<int64_t> subset_table(int nOrb, int maxN, int& maxSum) {
  maxSum = 0;
  for (int i = 0; i < maxN; ++i) maxSum += nOrb - 1 - i;
  std::vector<int64_t> t((size_t)(nOrb + 1) * (maxN + 1) * (maxSum + 1), 0);
  auto idx = [&](int j, int n, int s) { return ((size_t)j * (maxN + 1) + n) * (maxSum + 1) + s; };
  t[idx(0, 0, 0)] = 1;
  for (int j = 0; j < nOrb; ++j)
    for (int n = 0; n <= maxN; ++n)
      for (int s = 0; s <= maxSum; ++s) {
        int64_t v = t[idx(j, n, s)];
        if (n >= 1 && s >= j) v += t[idx(j, n - 1, s - j)];
        t[idx(j + 1, n, s)] = v;
      }
  return t;
}

__global__ void ms_unrank_kernel(uint64_t* __restrict__ states, int64_t dim, int nbrOrbitals, int nbrFermions,
                                 int orbitalSum, const int64_t* __restrict__ T, int maxN, int maxSum) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dim) return;
  int64_t rem = i;
  int n = nbrFermions, s = orbitalSum;
  uint64_t state = 0;
  for (int j = nbrOrbitals - 1; j >= 0 && n > 0; --j) {
    int64_t c = 0;
    if (s >= j && s - j <= maxSum) c = T[((int64_t)j * (maxN + 1) + (n - 1)) * (maxSum + 1) + (s - j)];
    if (rem < c) {
      state |= (1ull << j);
      --n;
      s -= j;
    } else {
      rem -= c;
    }
  }
  states[i] = state;
}

extern "C" int dh_sector_set_create(dh_context* c, dh_sector_set** out, int N, int lzMax, const int* totalLz,
                                    int nbrSectors) {
  if (!c || !out || !totalLz || nbrSectors <= 0) return DH_E_ARG;
  *out = nullptr;
  if (c->closed) return dh_fail(c, DH_E_STATE, "context destroyed");
  if (N <= 0 || lzMax < 0 || N > lzMax + 1) return dh_fail(c, DH_E_ARG, "need 0 < nbr_fermions <= lz_max + 1");
  if (lzMax > 62) return dh_fail(c, DH_E_RANGE, "lz_max must be <= 62");
  hipSetDevice(c->device);
  const int nOrb = lzMax + 1;
  int maxSum;
  std::vector<int64_t> T = subset_table(nOrb, N, maxSum);
  auto Tat = [&](int j, int n, int s) -> int64_t {
    if (j < 0 || j > nOrb || n < 0 || n > N || s < 0 || s > maxSum) return 0;
    return T[((size_t)j * (N + 1) + n) * (maxSum + 1) + s];
  };
  dh_sector_set* S = new dh_sector_set();
  S->ctx = c;
  S->nbrFermions = N;
  S->lzMax = lzMax;
  S->nbrOrbitals = nOrb;
  S->nbrSectors = nbrSectors;
  S->offsets.assign(nbrSectors + 1, 0);
  for (int i = 0; i < nbrSectors; ++i) {
    int twice = totalLz[i] + N * lzMax;
    if (twice < 0 || (twice & 1)) {
      delete S;
      return dh_fail(c, DH_E_ARG, "a total_lz has the wrong parity or range");
    }
    int sum = twice >> 1;
    int64_t d = Tat(nOrb, N, sum);
    if (d <= 0 || d >= (1ll << 31)) {
      delete S;
      return dh_fail(c, DH_E_ARG, "a sector is empty or larger than 2^31");
    }
    S->totalLz.push_back(totalLz[i]);
    S->orbitalSum.push_back(sum);
    S->offsets[i + 1] = S->offsets[i] + d;
  }
  S->totalDim = S->offsets[nbrSectors];
  S->rankBits = dh_rank_split(nOrb);
  S->loSize = 1ll << S->rankBits;
  S->hiSize = 1ll << (nOrb - S->rankBits);
  // rank tables: one low table, one high table per sector (the sectors' high tables on host threads:
  // 133 sectors of 2^20 entries at 34 orbitals)
  std::vector<uint32_t> lo(S->loSize, 0), hi((size_t)S->hiSize * nbrSectors, 0);
  for (uint64_t l = 0; l < (uint64_t)S->loSize; ++l) {
    int n = 0, s = 0;
    int64_t acc = 0;
    for (uint64_t r = l; r; r &= r - 1) {
      int j = __builtin_ctzll(r);
      ++n;
      s += j;
      acc += Tat(j, n, s);
    }
    lo[l] = (uint32_t)acc;
  }
  dh_parallel_for(nbrSectors, [&](int sec) {
    for (uint64_t h = 0; h < (uint64_t)S->hiSize; ++h) {
      int popH = __builtin_popcountll(h), sumH = 0;
      for (uint64_t r = h; r; r &= r - 1) sumH += S->rankBits + __builtin_ctzll(r);
      int n = N - popH, s = S->orbitalSum[sec] - sumH;
      if (n < 0 || s < 0) continue;
      int64_t acc = 0;
      for (uint64_t r = h; r; r &= r - 1) {
        int j = S->rankBits + __builtin_ctzll(r);
        ++n;
        s += j;
        acc += Tat(j, n, s);
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
    hipLaunchKernelGGL(ms_unrank_kernel, dim3((unsigned)((d + 255) / 256)), dim3(256), 0, c->stream,
                       S->d_states + S->offsets[sec], d, nOrb, N, S->orbitalSum[sec], d_T, N, maxSum);
  }
  hipError_t e = hipStreamSynchronize(c->stream);
  hipFree(d_T);
  if (e != hipSuccess) {
    dh_sector_set_destroy(S);
    return dh_fail(c, DH_E_HIP, std::string("sector set: ") + hipGetErrorString(e));
  }
  *out = S;
  return DH_OK;
}

extern "C" int dh_sector_set_destroy(dh_sector_set* S) {
  if (!S) return DH_E_ARG;
  dh_context* c = S->ctx;
  hipStreamSynchronize(c->stream);
  for (void* p : {(void*)S->d_states, (void*)S->d_rankLo, (void*)S->d_rankHi, (void*)S->d_rankLo16, (void*)S->d_offsets,
                  (void*)S->d_factors, (void*)S->d_pairOffset, (void*)S->d_work, (void*)S->d_coef,
                  (void*)S->d_pairInfo, (void*)S->d_sumPairOff, (void*)S->d_sumPairs, (void*)S->d_sumFactorOff,
                  (void*)S->d_cfactors, (void*)S->d_packedFactors, (void*)S->d_sumBase, (void*)S->d_orbit,
                  (void*)S->d_reorder, (void*)S->d_lut, (void*)S->d_phase, (void*)S->d_freeMask,
                  (void*)S->d_freeShift, (void*)S->d_partnerPos})
    if (p) hipFree(p);
  delete S;
  return DH_OK;
}

extern "C" int dh_sector_set_offsets(const dh_sector_set* S, int64_t* offsets) {
  if (!S || !offsets) return DH_E_ARG;
  std::memcpy(offsets, S->offsets.data(), sizeof(int64_t) * (S->nbrSectors + 1));
  return DH_OK;
}

extern "C" int dh_sector_set_get_states(dh_sector_set* S, uint64_t* states) {
  if (!S || !states) return DH_E_ARG;
  dh_context* c = S->ctx;
  DH_HIP(c, hipMemcpyAsync(states, S->d_states, sizeof(uint64_t) * S->totalDim, hipMemcpyDeviceToHost, c->stream));
  DH_HIP(c, hipStreamSynchronize(c->stream));
  return DH_OK;
}

extern "C" int dh_sector_set_factors(dh_sector_set* S, const double* factors, const int* pairOffset, int nbrFactors,
                                     double shift) {
  if (!S || !factors || !pairOffset) return DH_E_ARG;
  dh_context* c = S->ctx;
  const int nOrb = S->nbrOrbitals, nPairs = nOrb * (nOrb - 1) / 2;
  int expect = 0;
  for (int m1 = 0; m1 < nOrb; ++m1)
    for (int m2 = 0; m2 < m1; ++m2) {
      if (pairOffset[m1 * (m1 - 1) / 2 + m2] != expect) return dh_fail(c, DH_E_ARG, "pair_offset layout");
      int sum = m1 + m2;
      expect += std::min(sum, nOrb - 1) - (sum / 2 + 1) + 1;
    }
  if (pairOffset[nPairs] != expect || nbrFactors != expect) return dh_fail(c, DH_E_ARG, "nbr_factors layout");
  std::vector<double> packed;
  std::vector<int> sumBase;
  int rc = dh_pack_factors_checked(c, nOrb, factors, pairOffset, packed, sumBase);
  if (rc) return rc;
  hipStreamSynchronize(c->stream);
  for (void** p : {(void**)&S->d_factors, (void**)&S->d_pairOffset, (void**)&S->d_packedFactors, (void**)&S->d_sumBase}) {
    if (*p) hipFree(*p);
    *p = nullptr;
  }
  if ((rc = dh_alloc(c, (void**)&S->d_factors, sizeof(double) * std::max(1, nbrFactors))) ||
      (rc = dh_alloc(c, (void**)&S->d_pairOffset, sizeof(int) * (nPairs + 1))) ||
      (rc = dh_alloc(c, (void**)&S->d_packedFactors, sizeof(double) * std::max<size_t>(1, packed.size()))) ||
      (rc = dh_alloc(c, (void**)&S->d_sumBase, sizeof(int) * sumBase.size())))
    return rc;
  DH_HIP(c, hipMemcpyAsync(S->d_factors, factors, sizeof(double) * nbrFactors, hipMemcpyHostToDevice, c->stream));
  DH_HIP(c, hipMemcpyAsync(S->d_pairOffset, pairOffset, sizeof(int) * (nPairs + 1), hipMemcpyHostToDevice, c->stream));
  DH_HIP(c, hipMemcpyAsync(S->d_packedFactors, packed.data(), sizeof(double) * packed.size(), hipMemcpyHostToDevice,
                           c->stream));
  DH_HIP(c, hipMemcpyAsync(S->d_sumBase, sumBase.data(), sizeof(int) * sumBase.size(), hipMemcpyHostToDevice,
                           c->stream));
  DH_HIP(c, hipStreamSynchronize(c->stream));
  S->nbrFactors = nbrFactors;
  S->nbrPairs = nPairs;
  S->nbrPackedFactors = (int)packed.size();
  S->nbrPairSums = (int)sumBase.size();
  S->shift = shift;
  return DH_OK;
}

extern "C" int dh_sector_set_launch(dh_sector_set* S, int lanesPerRow, int blockSize) {
  if (!S) return DH_E_ARG;
  if (lanesPerRow != 1 && lanesPerRow != 2 && lanesPerRow != 4 && lanesPerRow != 8)
    return dh_fail(S->ctx, DH_E_ARG, "lanes_per_row in {1,2,4,8}");
  if (blockSize < 64 || blockSize > 1024 || (blockSize % 64)) return dh_fail(S->ctx, DH_E_ARG, "block_size");
  S->lanesPerRow = lanesPerRow;
  S->blockSize = blockSize;
  return DH_OK;
}

extern "C" int dh_sector_set_row_mode(dh_sector_set* S, int mode) {
  if (!S) return DH_E_ARG;
  if (mode != DH_INDEX_RANK && mode != DH_INDEX_RANK_FREE && mode != DH_INDEX_RANK_RANGE)
    return dh_fail(S->ctx, DH_E_ARG, "row mode must be DH_INDEX_RANK, DH_INDEX_RANK_FREE or DH_INDEX_RANK_RANGE");
  if (S->kind == DH_SECTORS_TORUS_TRANSLATIONS && mode == DH_INDEX_RANK_FREE)
    return dh_fail(S->ctx, DH_E_ARG, "torus sets with magnetic translations have one row routine");
  S->rowMode = mode;
  return DH_OK;
}

extern "C" int dh_sector_set_blas_launch(dh_sector_set* S, int axpyGroupsPerCU, int64_t dotsRowsPerTile) {
  if (!S) return DH_E_ARG;
  if (axpyGroupsPerCU < 1 || axpyGroupsPerCU > 64) return dh_fail(S->ctx, DH_E_ARG, "axpy groups per CU in [1, 64]");
  if (dotsRowsPerTile < 0 || (dotsRowsPerTile && dotsRowsPerTile < 256))
    return dh_fail(S->ctx, DH_E_ARG, "dots rows per tile: 0 (automatic) or >= 256");
  S->axpyGroupsPerCU = axpyGroupsPerCU;
  S->dotsRowsPerTile = dotsRowsPerTile;
  return DH_OK;
}

// work descriptor of the active sectors for tiles of `rowsPerTile` rows:
// d_work[0..na) = sector ids, d_work[na..2na] = tile prefix
int dh_sectors_upload_work(dh_sector_set* S, const unsigned char* active, int64_t rowsPerTile, int* nActive,
                       int64_t* totalTiles) {
  dh_context* c = S->ctx;
  std::vector<int64_t> w;
  std::vector<int64_t> prefix(1, 0);
  for (int s = 0; s < S->nbrSectors; ++s) {
    if (active && !active[s]) continue;
    w.push_back(s);
    int64_t d = S->offsets[s + 1] - S->offsets[s];
    prefix.push_back(prefix.back() + (d + rowsPerTile - 1) / rowsPerTile);
  }
  int na = (int)w.size();
  w.insert(w.end(), prefix.begin(), prefix.end());
  size_t bytes = sizeof(int64_t) * w.size();
  if (bytes > S->workCap) {
    if (S->d_work) {
      DH_HIP(c, hipStreamSynchronize(c->stream));
      hipFree(S->d_work);
    }
    S->d_work = nullptr;
    int rc = dh_alloc(c, (void**)&S->d_work, std::max<size_t>(bytes, 4096));
    if (rc) return rc;
    S->workCap = std::max<size_t>(bytes, 4096);
  }
  int rc = dh_upload_async(c, S->d_work, w.data(), bytes);
  if (rc) return rc;
  *nActive = na;
  *totalTiles = prefix.back();
  return DH_OK;
}

struct SectorsArgs {
  const uint64_t* states;
  const int64_t* offsets;
  int nbrOrbitals, nbrFactors, nbrSums;  // packed factors and their per-sum bases
  const double* factors;
  const int* sumBase;
  const uint32_t* rankLo;
  const uint32_t* rankHi;
  const uint16_t* rankLo16;  // the low table in 16 bits, staged in LDS (LO_LDS)
  int rankBits;
  int64_t loSize, hiSize;
  double shift;
  const int64_t* work;  // active ids then prefix
  int nActive;
  int64_t totalTiles;
};

// tile -> position j in the active list (prefix in LDS)
__device__ __forceinline__ int tile_sector(const int64_t* sPrefix, int nActive, int64_t tile) {
  int lo = 0, hi = nActive;  // sPrefix[lo] <= tile < sPrefix[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (sPrefix[mid] <= tile)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// S32: the orbitals fit a 32-bit word, so the row routine runs on 32-bit state words (used for
// k <= 2).  Every sector has fewer than 2^31 states (the set refuses larger ones), so indices are
// 32-bit in both forms.
// FREE: the created pairs come from the free-pair masks (dh_sphere_row_free).
template <int K, int LPR, bool LO_LDS, bool S32, bool FREE>
__global__ void __launch_bounds__(1024) sectors_hv_kernel(SectorsArgs a, const double* __restrict__ x,
                                                          double* __restrict__ y, int add) {
  extern __shared__ double smem[];
  double* sF = smem;
  int* sBase = reinterpret_cast<int*>(sF + a.nbrFactors);
  int64_t* sWork = reinterpret_cast<int64_t*>(smem) + ((a.nbrFactors * 8 + a.nbrSums * 4 + 7) / 8);
  uint16_t* sLo = reinterpret_cast<uint16_t*>(sWork + 2 * a.nActive + 1);
  for (int i = threadIdx.x; i < a.nbrFactors; i += blockDim.x) sF[i] = a.factors[i];
  for (int i = threadIdx.x; i < a.nbrSums; i += blockDim.x) sBase[i] = a.sumBase[i];
  for (int i = threadIdx.x; i < 2 * a.nActive + 1; i += blockDim.x) sWork[i] = a.work[i];
  if constexpr (LO_LDS)
    for (int i = threadIdx.x; i < a.loSize; i += blockDim.x) sLo[i] = a.rankLo16[i];
  __syncthreads();
  // the low table: 16-bit copy in LDS, or the 32-bit table in global memory
  auto loAt = [&](uint32_t i) -> uint32_t {
    if constexpr (LO_LDS)
      return sLo[i];
    else
      return a.rankLo[i];
  };
  const int64_t* sPrefix = sWork + a.nActive;
  const int lane = threadIdx.x % LPR;
  const int64_t rowsPerTile = blockDim.x / LPR;
  for (int64_t tile = dh_xcd_block(blockIdx.x, gridDim.x); tile < a.totalTiles; tile += gridDim.x) {
    const int j = tile_sector(sPrefix, a.nActive, tile);
    const int sec = (int)sWork[j];
    const int64_t off = a.offsets[sec];
    const int64_t dim = a.offsets[sec + 1] - off;
    const uint32_t* hi = a.rankHi + (int64_t)sec * a.hiSize;
    const int64_t local = (tile - sPrefix[j]) * rowsPerTile + threadIdx.x / LPR;
    const bool active = local < dim;
    const int64_t g = off + (active ? local : 0);
    const uint64_t s = active ? a.states[g] : 0ull;
    double acc[K];
#pragma unroll
    for (int v = 0; v < K; ++v) acc[v] = 0.0;
    if constexpr (S32) {
      const uint32_t last = (uint32_t)dim - 1u, mask = (1u << a.rankBits) - 1u;
      const int bits = a.rankBits;
      // clamped so that a target outside the sector can never address outside the vectors
      auto find = [&](uint32_t t) -> uint32_t { return min(last - loAt(t & mask) - hi[t >> bits], last); };
      if constexpr (FREE)
        dh_sphere_row_free<K, LPR>((uint32_t)s, a.nbrOrbitals, sF, sBase, lane, find, x + off * K, acc);
      else
        dh_sphere_row<K, LPR>((uint32_t)s, a.nbrOrbitals, sF, sBase, lane, find, x + off * K, acc);
    } else {
      // rankBits = ceil(orbitals / 2) <= 32, so t >> bits fits 32 bits
      const uint32_t last = (uint32_t)dim - 1u;
      const uint32_t mask = a.rankBits >= 32 ? 0xffffffffu : (1u << a.rankBits) - 1u;
      const int bits = a.rankBits;
      auto find = [&](uint64_t t) -> uint32_t {
        return min(last - loAt((uint32_t)t & mask) - hi[(uint32_t)(t >> bits)], last);
      };
      if constexpr (FREE)
        dh_sphere_row_free<K, LPR>(s, a.nbrOrbitals, sF, sBase, lane, find, x + off * K, acc);
      else
        dh_sphere_row<K, LPR>(s, a.nbrOrbitals, sF, sBase, lane, find, x + off * K, acc);
    }
    if constexpr (LPR > 1) {
#pragma unroll
      for (int v = 0; v < K; ++v)
        for (int o = LPR / 2; o > 0; o >>= 1) acc[v] += __shfl_xor(acc[v], o, LPR);
    }
    if (active && lane == 0) {
      const double* xr = x + g * K;
      double* yr = y + g * K;
#pragma unroll
      for (int v = 0; v < K; ++v) yr[v] = (add ? yr[v] : 0.0) + acc[v] + a.shift * xr[v];
    }
  }
}

static SectorsArgs sectors_args(dh_sector_set* S, int nActive, int64_t totalTiles) {
  SectorsArgs a;
  a.states = S->d_states;
  a.offsets = S->d_offsets;
  a.nbrOrbitals = S->nbrOrbitals;
  a.nbrFactors = S->nbrPackedFactors;
  a.nbrSums = S->nbrPairSums;
  a.factors = S->d_packedFactors;
  a.sumBase = S->d_sumBase;
  a.rankLo = S->d_rankLo;
  a.rankHi = S->d_rankHi;
  a.rankLo16 = S->d_rankLo16;
  a.rankBits = S->rankBits;
  a.loSize = S->loSize;
  a.hiSize = S->hiSize;
  a.shift = S->shift;
  a.work = S->d_work;
  a.nActive = nActive;
  a.totalTiles = totalTiles;
  return a;
}

template <int K, int LPR>
static int sectors_hv_launch(dh_sector_set* S, const double* x, double* y, int add, const unsigned char* active) {
  dh_context* c = S->ctx;
  int64_t rowsPerTile = S->blockSize / LPR;
  int na;
  int64_t tiles;
  int rc = dh_sectors_upload_work(S, active, rowsPerTile, &na, &tiles);
  if (rc) return rc;
  if (tiles == 0) return DH_OK;
  size_t base = ((size_t)S->nbrPackedFactors * 8 + (size_t)S->nbrPairSums * 4 + 7) / 8 * 8 + 8 * (2 * na + 1);
  // the 16-bit low table in LDS when it exists and the workgroup's tables stay within 64 KiB
  size_t withLo = base + 2 * S->loSize;
  bool loLds = S->d_rankLo16 && withLo <= 64 * 1024;
  size_t lds = loLds ? withLo : base;
  if (lds > 160 * 1024) return dh_fail(c, DH_E_RANGE, "staged tables exceed LDS");
  SectorsArgs a = sectors_args(S, na, tiles);
  auto launch = [&](auto kern) -> int {
    if (lds > 64 * 1024)
      DH_HIP(c, hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const int perCU = dh_blocks_per_cu(c, kern, S->blockSize, lds);
    int64_t grid = std::max<int64_t>(1, std::min<int64_t>(tiles, (int64_t)c->nbrCU * perCU));
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(S->blockSize), lds, c->stream, a, x, y, add);
    DH_HIP(c, hipGetLastError());
    return DH_OK;
  };
  // 32-bit words for one or two vectors, where the integer work of the row routine dominates; for
  // more vectors the 64-bit form is faster on MI355X (its lower occupancy keeps more of the x
  // gathers in cache: 11.3 against 16.5 ms at k = 8 over the 91 sectors of N=10 2S=27)
  const bool s32 = K <= 2 && S->nbrOrbitals <= 32;
  // DH_INDEX_RANK: the free-pair masks for k <= 2, the range loop for k > 2 (the faster per k on
  // MI355X, as for a single space)
  if (S->rowMode == DH_INDEX_RANK_FREE || (S->rowMode == DH_INDEX_RANK && K <= 2)) {
    if (loLds)
      return s32 ? launch(sectors_hv_kernel<K, LPR, true, (K <= 2), true>)
                 : launch(sectors_hv_kernel<K, LPR, true, false, true>);
    return s32 ? launch(sectors_hv_kernel<K, LPR, false, (K <= 2), true>)
               : launch(sectors_hv_kernel<K, LPR, false, false, true>);
  }
  if (loLds)
    return s32 ? launch(sectors_hv_kernel<K, LPR, true, (K <= 2), false>)
               : launch(sectors_hv_kernel<K, LPR, true, false, false>);
  return s32 ? launch(sectors_hv_kernel<K, LPR, false, (K <= 2), false>)
             : launch(sectors_hv_kernel<K, LPR, false, false, false>);
}

template <int K>
static int sectors_hv_lpr(dh_sector_set* S, const double* x, double* y, int add, const unsigned char* active) {
  switch (S->lanesPerRow) {
    case 1: return sectors_hv_launch<K, 1>(S, x, y, add, active);
    case 2: return sectors_hv_launch<K, 2>(S, x, y, add, active);
    case 4: return sectors_hv_launch<K, 4>(S, x, y, add, active);
    default: return sectors_hv_launch<K, 8>(S, x, y, add, active);
  }
}

extern "C" int dh_sectors_hv(dh_sector_set* S, const double* x, double* y, int k, int add,
                             const unsigned char* active) {
  if (!S || !x || !y) return DH_E_ARG;
  if (S->kind == DH_SECTORS_MOMENTUM) return dh_lattice_sectors_hv(S, x, y, k, add, active);
  if (S->kind == DH_SECTORS_TORUS_TRANSLATIONS) return dh_torus_translation_sectors_hv(S, x, y, k, add, active);
  if (!S->d_factors) return dh_fail(S->ctx, DH_E_STATE, "factors not set");
  switch (k) {
    case 1: return sectors_hv_lpr<1>(S, x, y, add, active);
    case 2: return sectors_hv_lpr<2>(S, x, y, add, active);
    case 4: return sectors_hv_lpr<4>(S, x, y, add, active);
    case 8: return sectors_hv_lpr<8>(S, x, y, add, active);
    case 16: return sectors_hv_lpr<16>(S, x, y, add, active);
    default: return dh_fail(S->ctx, DH_E_ARG, "k must be 1, 2, 4, 8 or 16");
  }
}

// ----------------------------------------------------------------------------- segmented BLAS
// dots: partial[tile][b*KK + i] over the tile's rows, reduced per active sector afterwards
template <int K, int G>
__global__ void __launch_bounds__(256) sectors_dots_kernel(const double* __restrict__ V, int nbrBlocks,
                                                           int64_t totalDim, const double* __restrict__ W,
                                                           const int64_t* __restrict__ offsets,
                                                           const int64_t* __restrict__ work, int nActive,
                                                           int64_t rowsPerTile, double* __restrict__ partial,
                                                           bool pairs) {
  constexpr int KK = K * K;
  __shared__ double red[256 / 64][G * KK];  // dh_dots_tile's per-wavefront sums
  const int64_t tile = blockIdx.y;  // blocks of one tile are adjacent: its W rows stay in L2
  const int64_t* prefix = work + nActive;
  int lo = 0, hi = nActive;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (prefix[mid] <= tile) lo = mid; else hi = mid;
  }
  const int sec = (int)work[lo];
  const int64_t off = offsets[sec], dim = offsets[sec + 1] - off;
  const int64_t r0 = (tile - prefix[lo]) * rowsPerTile;
  const int64_t r1 = min(dim, r0 + rowsPerTile);
  const int b0 = blockIdx.x * G;
  const int64_t total = (int64_t)nbrBlocks * KK;
  dh_dots_tile<K, G>(V, totalDim * K, nbrBlocks, b0, W, off + r0, off + r1, &red[0][0],
                     partial + tile * total + (int64_t)b0 * KK, pairs);
}

__global__ void sectors_reduce_kernel(const double* __restrict__ partial, const int64_t* __restrict__ work,
                                      int nActive, int64_t total, double* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= total) return;
  const int64_t* prefix = work + nActive;
  double s = 0.0;
  for (int64_t t = prefix[j]; t < prefix[j + 1]; ++t) s += partial[t * total + i];
  out[(int64_t)j * total + i] = s;
}

template <int K>
static int sectors_dots(dh_sector_set* S, const double* V, int nb, const double* W, const unsigned char* active,
                        double* h_out) {
  dh_context* c = S->ctx;
  constexpr int G = DotsGroup<K>::value;
  // automatic tile: at least 8192 rows (4096 for k >= 4, whose one-block workgroups are few:
  // 3.9-4.0 against 2.5-3.6 TB/s at k = 8 over the 91 sectors of N=10 2S=27), and no more
  // than about 32768 tiles
  const int64_t rowsPerTile = S->dotsRowsPerTile
                                  ? S->dotsRowsPerTile
                                  : std::max<int64_t>(K >= 4 ? 4096 : 8192, (S->totalDim + 32767) / 32768);
  int na;
  int64_t tiles;
  int rc = dh_sectors_upload_work(S, active, rowsPerTile, &na, &tiles);
  if (rc) return rc;
  if (tiles == 0) return DH_OK;
  const int64_t total = (int64_t)nb * K * K;
  double* scratch;
  if ((rc = dh_scratch(c, sizeof(double) * (size_t)(tiles * total + (int64_t)na * total), &scratch))) return rc;
  double* d_out = scratch + tiles * total;
  dim3 grid((unsigned)((nb + G - 1) / G), (unsigned)tiles);
  hipLaunchKernelGGL((sectors_dots_kernel<K, G>), grid, dim3(256), 0, c->stream, V, nb, S->totalDim, W, S->d_offsets,
                     S->d_work, na, rowsPerTile, scratch, K == 1 && dh_pairs_aligned(V, S->totalDim, W));
  hipLaunchKernelGGL(sectors_reduce_kernel, dim3((unsigned)((total + 255) / 256), (unsigned)na), dim3(256), 0,
                     c->stream, scratch, S->d_work, na, total, d_out);
  DH_HIP(c, hipGetLastError());
  double* pin;
  if ((rc = dh_pinned(c, sizeof(double) * na * total, &pin))) return rc;
  DH_HIP(c, hipMemcpyAsync(pin, d_out, sizeof(double) * na * total, hipMemcpyDeviceToHost, c->stream));
  DH_HIP(c, hipStreamSynchronize(c->stream));
  std::memcpy(h_out, pin, sizeof(double) * na * total);
  return DH_OK;
}

extern "C" int dh_sectors_block_dots(dh_sector_set* S, const double* V, int nb, const double* W, int k,
                                     const unsigned char* active, double* out) {
  if (!S || !V || !W || !out || nb <= 0) return DH_E_ARG;
  switch (k) {
    case 1: return sectors_dots<1>(S, V, nb, W, active, out);
    case 2: return sectors_dots<2>(S, V, nb, W, active, out);
    case 4: return sectors_dots<4>(S, V, nb, W, active, out);
    case 8: return sectors_dots<8>(S, V, nb, W, active, out);
    case 16: return sectors_dots<16>(S, V, nb, W, active, out);
    default: return dh_fail(S->ctx, DH_E_ARG, "k must be 1, 2, 4, 8 or 16");
  }
}

// axpy / rmul: W[row][q] (+)= sum over (b, p) of V_b[row][p] C_j[(b*K+p)*K+q]; rmul uses W itself.
// The coefficients sit in LDS in slabs of `slab` basis blocks (dynamic LDS of slab * K * K
// doubles, so short bases leave room for more workgroups per CU). Basis rows are read U blocks at
// a time, so one lane keeps U independent loads in flight; at k = 1 with the nontemporal hint (each
// basis row is read once per pass).
template <int K, bool RMUL>
__global__ void __launch_bounds__(256) sectors_axpy_kernel(const double* __restrict__ V, int nb, int64_t totalDim,
                                                           double* __restrict__ W,
                                                           const int64_t* __restrict__ offsets,
                                                           const int64_t* __restrict__ work, int nActive,
                                                           int64_t totalTiles, int slab, const double* __restrict__ C) {
  constexpr int KK = K * K;
  constexpr int U = K == 1 ? 8 : (K == 2 ? 4 : (K == 4 ? 2 : 1));
  extern __shared__ double sC[];
  const int64_t* prefix = work + nActive;
  const int64_t stride = totalDim * K;
  // each workgroup walks a contiguous range of tiles, so it stays in one sector for many tiles:
  // the sector is found by advancing through the prefix, and a coefficient matrix that fits one
  // LDS slab is loaded only when the sector changes
  const int64_t first = totalTiles * blockIdx.x / gridDim.x, last = totalTiles * (blockIdx.x + 1) / gridDim.x;
  int lo = 0;
  while (lo + 1 < nActive && prefix[lo + 1] <= first) ++lo;
  int staged = -1;
  for (int64_t tile = first; tile < last; ++tile) {
    while (prefix[lo + 1] <= tile) ++lo;
    const int sec = (int)work[lo];
    const int64_t off = offsets[sec], dim = offsets[sec + 1] - off;
    const int64_t r = (tile - prefix[lo]) * blockDim.x + threadIdx.x;
    const bool active = r < dim;
    const int64_t row = off + (active ? r : 0);
    const double* Cj = C + (int64_t)lo * nb * KK;
    double acc[K];
#pragma unroll
    for (int q = 0; q < K; ++q) acc[q] = 0.0;
    for (int s0 = 0; s0 < nb; s0 += slab) {
      const int ns = min(slab, nb - s0);
      if (nb > slab || staged != lo) {
        __syncthreads();
        for (int i = threadIdx.x; i < ns * KK; i += blockDim.x) sC[i] = Cj[(int64_t)s0 * KK + i];
        __syncthreads();
        staged = lo;
      }
      if (active) {
        const double* vb = (RMUL ? W : V + (int64_t)s0 * stride) + row * K;
        int b = 0;
        for (; b + U <= ns; b += U) {
          double v[U][K];
#pragma unroll
          for (int u = 0; u < U; ++u)
#pragma unroll
            for (int p = 0; p < K; ++p) v[u][p] = (RMUL || K > 1) ? vb[(int64_t)(b + u) * stride + p]
                                                                  : dh_stream_load(vb + (int64_t)(b + u) * stride + p);
#pragma unroll
          for (int u = 0; u < U; ++u)
#pragma unroll
            for (int p = 0; p < K; ++p)
#pragma unroll
              for (int q = 0; q < K; ++q) acc[q] += v[u][p] * sC[(b + u) * KK + p * K + q];
        }
        for (; b < ns; ++b) {
#pragma unroll
          for (int p = 0; p < K; ++p) {
            const double v = (RMUL || K > 1) ? vb[(int64_t)b * stride + p] : dh_stream_load(vb + (int64_t)b * stride + p);
#pragma unroll
            for (int q = 0; q < K; ++q) acc[q] += v * sC[b * KK + p * K + q];
          }
        }
      }
    }
    if (active) {
#pragma unroll
      for (int q = 0; q < K; ++q) {
        if (RMUL)
          W[row * K + q] = acc[q];
        else
          W[row * K + q] += acc[q];
      }
    }
  }
}

template <int K, bool RMUL>
static int sectors_axpy(dh_sector_set* S, const double* V, int nb, double* W, const unsigned char* active,
                        const double* h_c) {
  dh_context* c = S->ctx;
  int na;
  int64_t tiles;
  int rc = dh_sectors_upload_work(S, active, 256, &na, &tiles);
  if (rc) return rc;
  if (tiles == 0) return DH_OK;
  size_t bytes = sizeof(double) * (size_t)na * nb * K * K;
  if (bytes > S->coefCap) {
    if (S->d_coef) {
      DH_HIP(c, hipStreamSynchronize(c->stream));
      hipFree(S->d_coef);
    }
    S->d_coef = nullptr;
    if ((rc = dh_alloc(c, (void**)&S->d_coef, bytes))) return rc;
    S->coefCap = bytes;
  }
  if ((rc = dh_upload_async(c, S->d_coef, h_c, bytes))) return rc;
  // LDS slab: the whole coefficient block when it fits 32 KiB, else 32 KiB pieces
  const int slab = std::min(nb, std::max(1, 4096 / (K * K)));
  const size_t lds = sizeof(double) * (size_t)slab * K * K;
  int64_t grid = std::min<int64_t>(tiles, (int64_t)c->nbrCU * S->axpyGroupsPerCU);
  hipLaunchKernelGGL((sectors_axpy_kernel<K, RMUL>), dim3((unsigned)grid), dim3(256), lds, c->stream, V, nb,
                     S->totalDim, W, S->d_offsets, S->d_work, na, tiles, slab, S->d_coef);
  DH_HIP(c, hipGetLastError());
  return DH_OK;
}

extern "C" int dh_sectors_block_axpy(dh_sector_set* S, const double* V, int nb, double* W, int k,
                                     const unsigned char* active, const double* coef) {
  if (!S || !V || !W || !coef || nb <= 0) return DH_E_ARG;
  switch (k) {
    case 1: return sectors_axpy<1, false>(S, V, nb, W, active, coef);
    case 2: return sectors_axpy<2, false>(S, V, nb, W, active, coef);
    case 4: return sectors_axpy<4, false>(S, V, nb, W, active, coef);
    case 8: return sectors_axpy<8, false>(S, V, nb, W, active, coef);
    case 16: return sectors_axpy<16, false>(S, V, nb, W, active, coef);
    default: return dh_fail(S->ctx, DH_E_ARG, "k must be 1, 2, 4, 8 or 16");
  }
}

extern "C" int dh_sectors_block_rmul(dh_sector_set* S, double* W, int k, const unsigned char* active,
                                     const double* m) {
  if (!S || !W || !m) return DH_E_ARG;
  switch (k) {
    case 1: return sectors_axpy<1, true>(S, W, 1, W, active, m);
    case 2: return sectors_axpy<2, true>(S, W, 1, W, active, m);
    case 4: return sectors_axpy<4, true>(S, W, 1, W, active, m);
    case 8: return sectors_axpy<8, true>(S, W, 1, W, active, m);
    case 16: return sectors_axpy<16, true>(S, W, 1, W, active, m);
    default: return dh_fail(S->ctx, DH_E_ARG, "k must be 1, 2, 4, 8 or 16");
  }
}

extern "C" int64_t dh_sector_set_dimension(const dh_sector_set* S) { return S ? S->totalDim : -1; }
