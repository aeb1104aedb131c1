// On-the-fly Hamiltonian-vector products (gather form) for gfx950.
//
// Reference loops replaced:
//   fermions on the sphere: AbstractQHEOnSphereHamiltonian::LowLevelAddMultiply / HermitianLowLevelAddMultiply /
//     LowLevelMultipleAddMultiply (FQHE/src/Hamiltonian/AbstractQHEOnSphereHamiltonian.cc:410, :2563, :1217),
//     with FermionOnSphere::AA (:790), AdAd (:949) and FindStateIndex (:1431) inlined;
//   spin-1/2 chain: SpinChainHamiltonian::LowLevelAddMultiply (Spin/src/Hamiltonian/SpinChainHamiltonian.cc:257)
//     with Spin1_2Chain::SmiSpj (:655) and SziSzj (:561).
//
// The reference scatters (y[target] += H[target,i] x[i]).  These kernels gather instead:
// H is symmetric, so row r of H is the list of terms generated from state r, and
//   y[r] = sum_terms coef * x[target] + shift * x[r]
// is written once per destination row, with no atomics.  One row per group of
// `LPR` lanes (LPR = 1: one destination row per lane); the state word stays in
// registers, fermionic signs come from popcounts, factors are staged in LDS.
#include <algorithm>
#include <cstring>

#include "dh_index.h"
#include "dh_internal.h"
#include "dh_sphere_row.h"

// index lookup of the product kernels: where the rank tables live (both in LDS; the 16-bit low
// table in LDS and the high table in global memory; both in global memory), and the row routine
// (_X4: the created-pair loop unrolled four-wide; _FREE: only the free created pairs,
// dh_sphere_row_free)
enum {
  MODE_BSEARCH = 0,
  MODE_RANK_LDS = 1,
  MODE_RANK_GLOBAL = 2,
  MODE_RANK_LDS_X4 = 3,
  MODE_RANK_GLOBAL_X4 = 4,
  MODE_RANK_LDS_FREE = 5,
  MODE_RANK_GLOBAL_FREE = 6,
  MODE_RANK_LO16 = 7,
  MODE_RANK_LO16_X4 = 8,
  MODE_RANK_LO16_FREE = 9
};

template <int MODE>
struct ModeBase {
  static constexpr int value =
      MODE >= MODE_RANK_LO16 ? MODE_RANK_LO16
                             : (MODE >= MODE_RANK_LDS_FREE ? MODE - 4 : (MODE >= MODE_RANK_LDS_X4 ? MODE - 2 : MODE));
};

enum { ROW_RANGE = 0, ROW_X4 = 1, ROW_FREE = 2 };
template <int MODE>
struct ModeRow {
  static constexpr int value =
      (MODE == MODE_RANK_LDS_X4 || MODE == MODE_RANK_GLOBAL_X4 || MODE == MODE_RANK_LO16_X4) ? ROW_X4
      : (MODE == MODE_RANK_LDS_FREE || MODE == MODE_RANK_GLOBAL_FREE || MODE == MODE_RANK_LO16_FREE) ? ROW_FREE
                                                                                                   : ROW_RANGE;
};


struct SphereArgs {
  const uint64_t* states;
  int64_t dim;
  int nbrOrbitals, nbrFactors, nbrSums;  // packed factors and their per-sum bases
  const double* factors;
  const int* sumBase;
  const int* lut;
  int lutShift, lutSize, lutStride;  // staged LUT entry i = lut[i * lutStride]
  const uint32_t* rankLo;
  const uint32_t* rankHi;
  const uint16_t* rankLo16;
  int rankBits, rankLoSize, rankHiSize;
  double shift;
};

// ----------------------------------------------------------------------------- sphere H.v
template <int K>
__device__ __forceinline__ void gather_fma(double* acc, const double* __restrict__ x, int64_t idx, double coef) {
  const double* xr = x + idx * K;
  if constexpr (K % 2 == 0) {
#pragma unroll
    for (int v = 0; v < K; v += 2) {
      double2 p = *reinterpret_cast<const double2*>(xr + v);
      acc[v] += coef * p.x;
      acc[v + 1] += coef * p.y;
    }
  } else {
#pragma unroll
    for (int v = 0; v < K; ++v) acc[v] += coef * xr[v];
  }
}

template <int MODE>
__device__ __forceinline__ void sphere_stage(const SphereArgs& a, double*& sF, int*& sBase, int*& sLut,
                                             uint32_t*& sLo, uint32_t*& sHi, uint16_t*& sLo16) {
  extern __shared__ double smem[];
  sF = smem;
  sBase = reinterpret_cast<int*>(sF + a.nbrFactors);
  int* tail = sBase + a.nbrSums;
  for (int i = threadIdx.x; i < a.nbrFactors; i += blockDim.x) sF[i] = a.factors[i];
  for (int i = threadIdx.x; i < a.nbrSums; i += blockDim.x) sBase[i] = a.sumBase[i];
  if constexpr (MODE == MODE_BSEARCH) {
    sLut = tail;
    for (int i = threadIdx.x; i < a.lutSize; i += blockDim.x) sLut[i] = a.lut[(int64_t)i * a.lutStride];
  } else if constexpr (MODE == MODE_RANK_LDS) {
    sLo = reinterpret_cast<uint32_t*>(tail);
    sHi = sLo + a.rankLoSize;
    for (int i = threadIdx.x; i < a.rankLoSize; i += blockDim.x) sLo[i] = a.rankLo[i];
    for (int i = threadIdx.x; i < a.rankHiSize; i += blockDim.x) sHi[i] = a.rankHi[i];
  } else if constexpr (MODE == MODE_RANK_LO16) {
    sLo16 = reinterpret_cast<uint16_t*>(tail);
    sHi = const_cast<uint32_t*>(a.rankHi);
    for (int i = threadIdx.x; i < a.rankLoSize; i += blockDim.x) sLo16[i] = a.rankLo16[i];
  } else {
    sLo = const_cast<uint32_t*>(a.rankLo);
    sHi = const_cast<uint32_t*>(a.rankHi);
  }
  __syncthreads();
}

template <int MODE>
__device__ __forceinline__ int64_t sphere_find(const SphereArgs& a, uint64_t tgt, const int* sLut,
                                               const uint32_t* sLo, const uint32_t* sHi, const uint16_t* sLo16) {
  if constexpr (MODE == MODE_BSEARCH)
    return dh_find_bsearch(tgt, sLut, a.lutShift, a.states);
  else if constexpr (MODE == MODE_RANK_LO16)
    return dh_find_rank(tgt, sLo16, sHi, a.rankBits, a.dim);
  else
    return dh_find_rank(tgt, sLo, sHi, a.rankBits, a.dim);
}

// WORDS (rank modes): 1 = the orbitals fit a 32-bit word and the space has fewer than 2^31 states,
// so the row routine runs on 32-bit state words and 32-bit indices (used for k <= 2); 2 = 64-bit
// state words with 32-bit indices (fewer than 2^31 states: the rank, its clamp and the gather
// offset in 32-bit arithmetic); 0 = 64-bit words and indices.
template <int K, int LPR, int MODE, int WORDS>
__global__ void __launch_bounds__(1024) sphere_hv_kernel(SphereArgs a, const double* __restrict__ x,
                                                         double* __restrict__ y, int add, int64_t first,
                                                         int64_t count) {
  constexpr int BASE = ModeBase<MODE>::value;
  double* sF;
  int *sBase, *sLut = nullptr;
  uint32_t *sLo = nullptr, *sHi = nullptr;
  uint16_t* sLo16 = nullptr;
  sphere_stage<BASE>(a, sF, sBase, sLut, sLo, sHi, sLo16);
  constexpr int ROW = ModeRow<MODE>::value;
  // the low table entry: 16-bit copy in LDS (MODE_RANK_LO16), else the 32-bit table (LDS or global)
  auto loAt = [&](uint32_t i) -> uint32_t {
    if constexpr (BASE == MODE_RANK_LO16)
      return sLo16[i];
    else
      return sLo[i];
  };
  const int lane = threadIdx.x % LPR;
  const int64_t rowsPerBlock = blockDim.x / LPR;
  const int64_t nbrTiles = (count + rowsPerBlock - 1) / rowsPerBlock;
  const unsigned vb = dh_xcd_block(blockIdx.x, gridDim.x);
  for (int64_t tile = vb; tile < nbrTiles; tile += gridDim.x) {
    const int64_t local = tile * rowsPerBlock + threadIdx.x / LPR;
    const bool active = local < count;
    const int64_t row = first + (active ? local : 0);
    const uint64_t s = active ? a.states[row] : 0ull;
    double acc[K];
#pragma unroll
    for (int v = 0; v < K; ++v) acc[v] = 0.0;
    if constexpr (WORDS == 1 && BASE != MODE_BSEARCH) {
      const uint32_t last = (uint32_t)a.dim - 1u, mask = (1u << a.rankBits) - 1u;
      const int bits = a.rankBits;
      // clamped so that a target outside the space can never address outside the vectors
      auto find = [&](uint32_t t) -> uint32_t { return min(last - loAt(t & mask) - sHi[t >> bits], last); };
      if constexpr (ROW == ROW_FREE)
        dh_sphere_row_free<K, LPR>((uint32_t)s, a.nbrOrbitals, sF, sBase, lane, find, x, acc);
      else if constexpr (ROW == ROW_X4)
        dh_sphere_row_x4<K, LPR>((uint32_t)s, a.nbrOrbitals, sF, sBase, lane, find, x, acc);
      else
        dh_sphere_row<K, LPR>((uint32_t)s, a.nbrOrbitals, sF, sBase, lane, find, x, acc);
    } else if constexpr (WORDS == 2 && BASE != MODE_BSEARCH) {
      // rankBits = ceil(orbitals / 2) <= 32, so t >> bits fits 32 bits
      const uint32_t last = (uint32_t)a.dim - 1u;
      const uint32_t mask = a.rankBits >= 32 ? 0xffffffffu : (1u << a.rankBits) - 1u;
      const int bits = a.rankBits;
      auto find = [&](uint64_t t) -> uint32_t {
        return min(last - loAt((uint32_t)t & mask) - sHi[(uint32_t)(t >> bits)], last);
      };
      if constexpr (ROW == ROW_FREE)
        dh_sphere_row_free<K, LPR>(s, a.nbrOrbitals, sF, sBase, lane, find, x, acc);
      else if constexpr (ROW == ROW_X4)
        dh_sphere_row_x4<K, LPR>(s, a.nbrOrbitals, sF, sBase, lane, find, x, acc);
      else
        dh_sphere_row<K, LPR>(s, a.nbrOrbitals, sF, sBase, lane, find, x, acc);
    } else {
      auto find = [&](uint64_t t) -> int64_t { return sphere_find<BASE>(a, t, sLut, sLo, sHi, sLo16); };
      if constexpr (ROW == ROW_FREE)
        dh_sphere_row_free<K, LPR>(s, a.nbrOrbitals, sF, sBase, lane, find, x, acc);
      else if constexpr (ROW == ROW_X4)
        dh_sphere_row_x4<K, LPR>(s, a.nbrOrbitals, sF, sBase, lane, find, x, acc);
      else
        dh_sphere_row<K, LPR>(s, a.nbrOrbitals, sF, sBase, lane, find, x, acc);
    }
    if constexpr (LPR > 1) {
#pragma unroll
      for (int v = 0; v < K; ++v)
        for (int o = LPR / 2; o > 0; o >>= 1) acc[v] += __shfl_xor(acc[v], o, LPR);
    }
    if (active && lane == 0) {
      const double* xr = x + row * K;
      double* yr = y + row * K;
#pragma unroll
      for (int v = 0; v < K; ++v) yr[v] = (add ? yr[v] : 0.0) + acc[v] + a.shift * xr[v];
    }
  }
}

// ----------------------------------------------------------------------------- spin H.v
struct SpinArgs {
  const uint32_t* words;
  int64_t dim;
  int L, periodic;
  const uint32_t* binom;
  const int* lut;
  int lutShift, lutSize;
  double shift;
  double jzSum;  // sum of Jz over the bonds of the chain
  double halfJ[32], Jz[32];
};

// Only antiparallel bonds contribute off-diagonal terms (and lower the diagonal by Jz/2 each):
// they are the set bits of s ^ (s >> 1).  Their target indices are computed G at a time before
// the G gathers of x are issued, so each lane keeps G independent loads in flight instead of
// waiting on one gather per bond.  This kernel finds targets by the LUT-seeded binary search of
// Spin1_2Chain::FindStateIndex; spin_rank_kernel below computes them from binomials.
template <int K>
__global__ void __launch_bounds__(1024) spin_bsearch_kernel(SpinArgs a, const double* __restrict__ x,
                                                            double* __restrict__ y, int add, int64_t first,
                                                            int64_t count) {
  constexpr int G = K == 1 ? 8 : (K == 2 ? 4 : 2);
  __shared__ double sHalfJ[32], sJz[32];
  extern __shared__ int sLut[];
  if (threadIdx.x < 32) {
    sHalfJ[threadIdx.x] = a.halfJ[threadIdx.x];
    sJz[threadIdx.x] = a.Jz[threadIdx.x];
  }
  for (int i = threadIdx.x; i < a.lutSize; i += blockDim.x) sLut[i] = a.lut[i];
  __syncthreads();
  const int L = a.L;
  const uint32_t mask = (L == 32) ? 0xffffffffu : ((1u << L) - 1u);
  const uint32_t openBonds = (1u << (L - 1)) - 1u;  // bonds (i, i+1), i < L-1
  const int64_t nbrTiles = (count + blockDim.x - 1) / blockDim.x;
  const unsigned vb = dh_xcd_block(blockIdx.x, gridDim.x);
  for (int64_t tile = vb; tile < nbrTiles; tile += gridDim.x) {
    const int64_t local = tile * blockDim.x + threadIdx.x;
    if (local >= count) continue;
    const int64_t row = first + local;
    const uint32_t s = a.words[row] & mask;
    uint32_t anti = (s ^ (s >> 1)) & openBonds;
    double diag = a.shift + 0.25 * a.jzSum;
    for (uint32_t r = anti; r; r &= r - 1) diag -= 0.5 * sJz[__builtin_ctz(r)];
    const bool wrap = a.periodic && (((s >> (L - 1)) ^ s) & 1u);
    if (wrap) diag -= 0.5 * sJz[L - 1];
    double acc[K];
    {
      const double* xr = x + row * K;
#pragma unroll
      for (int v = 0; v < K; ++v) acc[v] = diag * xr[v];
    }
    while (anti) {
      int64_t idx[G];
      double coef[G];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        idx[g] = row;  // padding entries gather the row itself with a zero coefficient
        coef[g] = 0.0;
        if (anti) {
          const int i = __builtin_ctz(anti);
          anti &= anti - 1;
          idx[g] = dh_spin_find_bsearch(s ^ (3u << i), L, sLut, a.lutShift, a.words);
          coef[g] = sHalfJ[i];
        }
      }
      double v[G][K];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const double* xr = x + idx[g] * K;
        if constexpr (K % 2 == 0) {
#pragma unroll
          for (int q = 0; q < K; q += 2) {
            double2 p = *reinterpret_cast<const double2*>(xr + q);
            v[g][q] = p.x;
            v[g][q + 1] = p.y;
          }
        } else {
#pragma unroll
          for (int q = 0; q < K; ++q) v[g][q] = xr[q];
        }
      }
#pragma unroll
      for (int g = 0; g < G; ++g)
#pragma unroll
        for (int q = 0; q < K; ++q) acc[q] += coef[g] * v[g][q];
    }
    if (wrap) {
      const uint32_t t = s ^ (1u | (1u << (L - 1)));
      gather_fma<K>(acc, x, dh_spin_find_bsearch(t, L, sLut, a.lutShift, a.words), sHalfJ[L - 1]);
    }
    double* yr = y + row * K;
#pragma unroll
    for (int v = 0; v < K; ++v) yr[v] = (add ? yr[v] : 0.0) + acc[v];
  }
}

// Rank-lookup spin product with 32-bit index arithmetic (a spin-1/2 chain of L <= 32 sites has
// at most C(32, 16) < 2^31 states in one Sz sector).  The binomials are staged in LDS in the two
// forms the product reads, with a power-of-two row stride so an entry is one shift-or away:
//   sStep[i * 64 + c] = C(L-2-i, c-1): index distance of the flip of bond (i, i+1) when c spins up
//                       sit at bits >= i;
//   sRank[i * 64 + c] = C(L-1-i, c): contribution of an up spin at bit i with c ups at bits >= i
//                       to the number of states ordered after the state (the wrap bond's rank).
// UNIFORM: every bond has the same J and Jz (the Heisenberg and XXZ chains): the diagonal is one
// popcount and the coefficient of every flip is J/2.
// Targets are gathered G at a time.  The lanes of a wavefront have different numbers of
// antiparallel bonds and the wavefront runs as many groups as its busiest lane, so a smaller G
// wastes fewer slots; for one vector G = 4 measured fastest (G = 8: 3-4 % slower, 12: 10 %).
template <int K, bool UNIFORM>
__global__ void __launch_bounds__(1024) spin_rank_kernel(SpinArgs a, const double* __restrict__ x,
                                                         double* __restrict__ y, int add, int64_t first,
                                                         int64_t count) {
  constexpr int G = K == 1 ? 4 : (K == 2 ? 4 : 2);
  __shared__ uint32_t sStep[32 * 64];
  __shared__ uint32_t sRank[32 * 64];
  __shared__ double sHalfJ[32], sJz[32];
  const int L = a.L;
  for (int e = threadIdx.x; e < 32 * 64; e += blockDim.x) {
    const int i = e >> 6, c = e & 63;
    uint32_t st = 0, rk = 0;
    if (i <= L - 2 && c >= 1 && c - 1 <= L - 2 - i) st = a.binom[(L - 2 - i) * 33 + (c - 1)];
    if (i <= L - 1 && c <= L - 1 - i) rk = a.binom[(L - 1 - i) * 33 + c];
    sStep[e] = st;
    sRank[e] = rk;
  }
  if (threadIdx.x < 32) {
    sHalfJ[threadIdx.x] = a.halfJ[threadIdx.x];
    sJz[threadIdx.x] = a.Jz[threadIdx.x];
  }
  __syncthreads();
  const uint32_t mask = (L == 32) ? 0xffffffffu : ((1u << L) - 1u);
  const uint32_t openBonds = (1u << (L - 1)) - 1u;  // bonds (i, i+1), i < L-1
  const uint32_t last = (uint32_t)a.dim - 1u;
  const double halfJ0 = a.halfJ[0], base = a.shift + 0.25 * a.jzSum, halfJz0 = 0.5 * a.Jz[0];
  const int64_t nbrTiles = (count + blockDim.x - 1) / blockDim.x;
  const unsigned vb = dh_xcd_block(blockIdx.x, gridDim.x);
  for (int64_t tile = vb; tile < nbrTiles; tile += gridDim.x) {
    const int64_t local = tile * blockDim.x + threadIdx.x;
    if (local >= count) continue;
    const uint32_t row = (uint32_t)(first + local);
    const uint32_t s = a.words[row] & mask;
    uint32_t anti = (s ^ (s >> 1)) & openBonds;
    const bool wrap = a.periodic && (((s >> (L - 1)) ^ s) & 1u);
    double diag;
    if constexpr (UNIFORM) {
      diag = base - halfJz0 * (double)(__popc(anti) + (wrap ? 1 : 0));
    } else {
      diag = base;
      for (uint32_t r = anti; r; r &= r - 1) diag -= 0.5 * sJz[__builtin_ctz(r)];
      if (wrap) diag -= 0.5 * sJz[L - 1];
    }
    double acc[K];
    {
      const double* xr = x + (size_t)row * K;
#pragma unroll
      for (int v = 0; v < K; ++v) acc[v] = diag * xr[v];
    }
    while (anti) {
      uint32_t idx[G];
      double coef[G];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        idx[g] = row;  // padding entries gather the row itself with a zero coefficient
        coef[g] = 0.0;
        if (anti) {
          const int i = __builtin_ctz(anti);
          anti &= anti - 1;
          const uint32_t d = sStep[(i << 6) | __popc(s >> i)];
          const uint32_t t = ((s >> i) & 1u) ? row + d : row - d;
          idx[g] = min(t, last);  // a state of the space never leaves [0, dim): malformed words stay in bounds
          coef[g] = UNIFORM ? halfJ0 : sHalfJ[i];
        }
      }
      double v[G][K];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const double* xr = x + (size_t)idx[g] * K;
        if constexpr (K % 2 == 0) {
#pragma unroll
          for (int q = 0; q < K; q += 2) {
            double2 p = *reinterpret_cast<const double2*>(xr + q);
            v[g][q] = p.x;
            v[g][q + 1] = p.y;
          }
        } else {
#pragma unroll
          for (int q = 0; q < K; ++q) v[g][q] = xr[q];
        }
      }
#pragma unroll
      for (int g = 0; g < G; ++g)
#pragma unroll
        for (int q = 0; q < K; ++q) acc[q] += coef[g] * v[g][q];
    }
    if (wrap) {
      const uint32_t t = s ^ (1u | (1u << (L - 1)));
      uint32_t after = 0;
      for (uint32_t r = t; r; r &= r - 1) {
        const int i = __builtin_ctz(r);
        after += sRank[(i << 6) | __popc(t >> i)];
      }
      gather_fma<K>(acc, x, (int64_t)min(last - after, last), UNIFORM ? halfJ0 : sHalfJ[L - 1]);
    }
    double* yr = y + (size_t)row * K;
#pragma unroll
    for (int v = 0; v < K; ++v) yr[v] = (add ? yr[v] : 0.0) + acc[v];
  }
}

// ----------------------------------------------------------------------------- host plumbing
static SphereArgs sphere_args(const dh_hamiltonian* h, int lutBitsUsed) {
  const dh_basis* b = h->basis;
  SphereArgs a;
  a.states = b->d_states;
  a.dim = b->dim;
  a.nbrOrbitals = b->nbrOrbitals;
  a.nbrFactors = h->nbrPackedFactors;
  a.nbrSums = h->nbrPairSums;
  a.factors = h->d_packedFactors;
  a.sumBase = h->d_sumBase;
  a.lut = b->d_lut;
  int drop = b->lutBits - lutBitsUsed;
  a.lutStride = 1 << drop;
  a.lutShift = b->lutShift + drop;
  a.lutSize = (1 << lutBitsUsed) + 1;
  a.rankLo = b->d_rankLo;
  a.rankHi = b->d_rankHi;
  a.rankLo16 = b->d_rankLo16;
  a.rankBits = b->rankBits;
  a.rankLoSize = (int)b->rankLoSize;
  a.rankHiSize = (int)b->rankHiSize;
  a.shift = h->shift;
  return a;
}

static const size_t kLdsLimit = 160 * 1024;

// LDS bytes of the staged tables for a given mode
static size_t sphere_lds(const dh_hamiltonian* h, int mode, int lutBitsUsed) {
  size_t bytes = sizeof(double) * h->nbrPackedFactors + sizeof(int) * h->nbrPairSums;
  bytes = (bytes + 7) & ~size_t(7);
  if (mode == MODE_BSEARCH) bytes += sizeof(int) * ((size_t(1) << lutBitsUsed) + 1);
  if (mode == MODE_RANK_LDS) bytes += sizeof(uint32_t) * (h->basis->rankLoSize + h->basis->rankHiSize);
  if (mode == MODE_RANK_LO16) bytes += sizeof(uint16_t) * h->basis->rankLoSize;
  return bytes;
}

// where the rank tables of a sphere product are read from (MODE_RANK_LDS, MODE_RANK_LO16 or
// MODE_RANK_GLOBAL)
static int sphere_rank_tables(const dh_hamiltonian* h) {
  if (h->ldsTables && sphere_lds(h, MODE_RANK_LDS, 0) <= 64 * 1024) return MODE_RANK_LDS;
  if (h->ldsTables && h->basis->d_rankLo16 && sphere_lds(h, MODE_RANK_LO16, 0) <= 96 * 1024) return MODE_RANK_LO16;
  return MODE_RANK_GLOBAL;
}

template <typename Kern>
static int launch_persistent(dh_context* c, Kern kern, int blockSize, size_t lds, int64_t nbrTiles, dim3& grid) {
  const int perCU = dh_blocks_per_cu(c, kern, blockSize, lds);
  int64_t g = std::min<int64_t>(nbrTiles, (int64_t)c->nbrCU * perCU);
  grid = dim3((unsigned)std::max<int64_t>(g, 1));
  return DH_OK;
}

template <int K, int LPR, int MODE>
static int sphere_launch(dh_hamiltonian* h, int bs, const double* x, double* y, int add, int64_t first,
                         int64_t count) {
  dh_context* c = h->ctx;
  int lutBitsUsed = std::min(h->basis->lutBits, 13);
  SphereArgs a = sphere_args(h, lutBitsUsed);
  size_t lds = sphere_lds(h, ModeBase<MODE>::value, lutBitsUsed);
  if (lds > kLdsLimit) return dh_fail(c, DH_E_RANGE, "staged tables exceed LDS");
  // 32-bit words for one or two vectors only, as for the sector sets; 32-bit indices whenever the
  // space has fewer than 2^31 states
  const bool idx32 = MODE != MODE_BSEARCH && h->basis->dim < (1ll << 31);
  const bool s32 = K <= 2 && idx32 && h->basis->nbrOrbitals <= 32;
  constexpr int W1 = (K <= 2 && MODE != MODE_BSEARCH) ? 1 : 0;
  constexpr int W2 = MODE != MODE_BSEARCH ? 2 : 0;
  auto kern = s32 ? sphere_hv_kernel<K, LPR, MODE, W1>
                  : (idx32 ? sphere_hv_kernel<K, LPR, MODE, W2> : sphere_hv_kernel<K, LPR, MODE, 0>);
  if (lds > 64 * 1024) DH_HIP(c, hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int64_t rowsPerBlock = bs / LPR;
  dim3 grid;
  launch_persistent(c, kern, bs, lds, (count + rowsPerBlock - 1) / rowsPerBlock, grid);
  hipLaunchKernelGGL(kern, grid, dim3(bs), lds, c->stream, a, x, y, add, first, count);
  DH_HIP(c, hipGetLastError());
  return DH_OK;
}

template <int K, int LPR>
static int sphere_dispatch_mode(dh_hamiltonian* h, int bs, const double* x, double* y, int add, int64_t first,
                                int64_t count) {
  if (h->indexMode == DH_INDEX_BSEARCH) return sphere_launch<K, LPR, MODE_BSEARCH>(h, bs, x, y, add, first, count);
  // where the rank tables go: both in LDS when they fit 64 KiB, else the 16-bit low table in LDS
  // (the high table from L2), else both from global memory
  const int where = sphere_rank_tables(h);
  const bool tablesInLds = where == MODE_RANK_LDS, lo16 = where == MODE_RANK_LO16;
  if (h->indexMode == DH_INDEX_RANK_X4) {
    if (tablesInLds) return sphere_launch<K, LPR, MODE_RANK_LDS_X4>(h, bs, x, y, add, first, count);
    if (lo16) return sphere_launch<K, LPR, MODE_RANK_LO16_X4>(h, bs, x, y, add, first, count);
    return sphere_launch<K, LPR, MODE_RANK_GLOBAL_X4>(h, bs, x, y, add, first, count);
  }
  // DH_INDEX_RANK: the free-pair masks for k <= 2 (fewer instructions per term), the range loop for
  // k > 2 (where the masks' extra registers cost more than they save)
  const bool freePairs = h->indexMode == DH_INDEX_RANK_FREE || (h->indexMode == DH_INDEX_RANK && K <= 2);
  if (freePairs) {
    if (tablesInLds) return sphere_launch<K, LPR, MODE_RANK_LDS_FREE>(h, bs, x, y, add, first, count);
    if (lo16) return sphere_launch<K, LPR, MODE_RANK_LO16_FREE>(h, bs, x, y, add, first, count);
    return sphere_launch<K, LPR, MODE_RANK_GLOBAL_FREE>(h, bs, x, y, add, first, count);
  }
  if (tablesInLds) return sphere_launch<K, LPR, MODE_RANK_LDS>(h, bs, x, y, add, first, count);
  if (lo16) return sphere_launch<K, LPR, MODE_RANK_LO16>(h, bs, x, y, add, first, count);
  return sphere_launch<K, LPR, MODE_RANK_GLOBAL>(h, bs, x, y, add, first, count);
}

template <int K>
static int sphere_dispatch_lpr(dh_hamiltonian* h, const double* x, double* y, int add, int64_t first,
                               int64_t count) {
  int lpr = h->lanesPerRow, bs = h->blockSize;
  if (h->autoLaunch && h->basis->dim >= (int64_t)h->ctx->nbrCU * 256) {
    // the fastest shapes measured on MI355X for N=12 2S=33 Lz=0 (7.8 M rows).  With the 16-bit low
    // table in LDS a workgroup holds ~60 KiB, two per CU: one lane per row with 1024-thread
    // workgroups for k <= 2 (6.1 ms at k = 1, against 10.6 ms at 256 threads), 512 for k > 2.
    // With both tables in global memory: two lanes per row, 256 threads for k <= 2, 1024 above.
    if (sphere_rank_tables(h) == MODE_RANK_LO16) {
      lpr = K == 2 ? 2 : 1;  // k = 2: two lanes per row (see below)
      bs = K <= 2 ? 1024 : 512;
    } else {
      lpr = 2;
      bs = K <= 2 ? 256 : 1024;
    }
  }
  // k = 2 with one lane per row faulted the GPU twice on N=5 2S=33 (free-pair rows, with the low
  // rank table in LDS and in global memory) with no cause found in the code; k = 2 runs with two
  // lanes per row, as tested, until one is
  if (K == 2 && lpr == 1) lpr = 2;
  switch (lpr) {
    case 1: return sphere_dispatch_mode<K, 1>(h, bs, x, y, add, first, count);
    case 2: return sphere_dispatch_mode<K, 2>(h, bs, x, y, add, first, count);
    case 4: return sphere_dispatch_mode<K, 4>(h, bs, x, y, add, first, count);
    case 8: return sphere_dispatch_mode<K, 8>(h, bs, x, y, add, first, count);
    default: return dh_fail(h->ctx, DH_E_ARG, "lanes_per_row must be 1, 2, 4 or 8 for this Hamiltonian");
  }
}

static SpinArgs spin_args(const dh_hamiltonian* h) {
  const dh_basis* b = h->basis;
  SpinArgs a;
  a.words = b->d_words;
  a.dim = b->dim;
  a.L = b->chainLength;
  a.periodic = h->periodic;
  a.binom = b->d_binom;
  a.lut = b->d_lut;
  a.lutShift = b->lutShift;
  a.lutSize = (1 << b->lutBits) + 1;
  a.shift = h->shift;
  a.jzSum = 0.0;
  for (int i = 0; i < 32; ++i) {
    a.halfJ[i] = 0.5 * h->J[i];
    a.Jz[i] = h->Jz[i];
    if (i < a.L - 1 || (i == a.L - 1 && a.periodic)) a.jzSum += h->Jz[i];
  }
  return a;
}

template <int K>
static int spin_bsearch_launch(dh_hamiltonian* h, const double* x, double* y, int add, int64_t first,
                               int64_t count) {
  dh_context* c = h->ctx;
  SpinArgs a = spin_args(h);
  size_t lds = sizeof(int) * (size_t)a.lutSize;
  auto kern = spin_bsearch_kernel<K>;
  dim3 grid;
  launch_persistent(c, kern, h->blockSize, lds, (count + h->blockSize - 1) / h->blockSize, grid);
  hipLaunchKernelGGL(kern, grid, dim3(h->blockSize), lds, c->stream, a, x, y, add, first, count);
  DH_HIP(c, hipGetLastError());
  return DH_OK;
}

template <int K, bool UNIFORM>
static int spin_rank_launch(dh_hamiltonian* h, const double* x, double* y, int add, int64_t first, int64_t count) {
  dh_context* c = h->ctx;
  SpinArgs a = spin_args(h);
  auto kern = spin_rank_kernel<K, UNIFORM>;
  dim3 grid;
  launch_persistent(c, kern, h->blockSize, 0, (count + h->blockSize - 1) / h->blockSize, grid);
  hipLaunchKernelGGL(kern, grid, dim3(h->blockSize), 0, c->stream, a, x, y, add, first, count);
  DH_HIP(c, hipGetLastError());
  return DH_OK;
}

// every bond of the chain (the L-1 open bonds, and the wrap bond when periodic) has the same J and Jz
static bool spin_uniform(const dh_hamiltonian* h) {
  const int L = h->basis->chainLength, nb = h->periodic ? L : L - 1;
  for (int i = 1; i < nb; ++i)
    if (h->J[i] != h->J[0] || h->Jz[i] != h->Jz[0]) return false;
  return true;
}

template <int K>
static int spin_dispatch(dh_hamiltonian* h, const double* x, double* y, int add, int64_t first, int64_t count) {
  if (h->indexMode == DH_INDEX_BSEARCH) return spin_bsearch_launch<K>(h, x, y, add, first, count);
  if (h->basis->dim >= (1ll << 31)) return dh_fail(h->ctx, DH_E_RANGE, "spin sector of 2^31 states or more");
  if (spin_uniform(h)) return spin_rank_launch<K, true>(h, x, y, add, first, count);
  return spin_rank_launch<K, false>(h, x, y, add, first, count);
}

extern "C" int dh_hv(dh_hamiltonian* h, const double* x, double* y, int k, int add, int64_t first, int64_t count) {
  if (!h || !x || !y) return DH_E_ARG;
  dh_context* c = h->ctx;
  int64_t dim = h->basis->dim;
  if (count < 0) count = dim - first;
  if (first < 0 || first + count > dim) return dh_fail(c, DH_E_ARG, "row range outside the space");
  if (count == 0) return DH_OK;
  if (h->d_csrRow) return dh_csr_hv(h, x, y, k, add, first, count);
  if (h->kind == DH_HAM_SPHERE) {
    if (!h->d_factors) return dh_fail(c, DH_E_STATE, "factors not set");
    switch (k) {
      case 1: return sphere_dispatch_lpr<1>(h, x, y, add, first, count);
      case 2: return sphere_dispatch_lpr<2>(h, x, y, add, first, count);
      case 4: return sphere_dispatch_lpr<4>(h, x, y, add, first, count);
      case 8: return sphere_dispatch_lpr<8>(h, x, y, add, first, count);
      case 16: return sphere_dispatch_lpr<16>(h, x, y, add, first, count);
      default: return dh_fail(c, DH_E_ARG, "k must be 1, 2, 4, 8 or 16");
    }
  }
  if (h->kind == DH_HAM_SPIN) {
    switch (k) {
      case 1: return spin_dispatch<1>(h, x, y, add, first, count);
      case 2: return spin_dispatch<2>(h, x, y, add, first, count);
      case 4: return spin_dispatch<4>(h, x, y, add, first, count);
      case 8: return spin_dispatch<8>(h, x, y, add, first, count);
      case 16: return spin_dispatch<16>(h, x, y, add, first, count);
      default: return dh_fail(c, DH_E_ARG, "k must be 1, 2, 4, 8 or 16");
    }
  }
  if (h->kind == DH_HAM_SPHERE_NBODY) return dh_nbody_hv(h, x, y, k, add, first, count);
  return dh_fail(c, DH_E_STATE, "unknown Hamiltonian kind");
}

// ----------------------------------------------------------------------------- k separate vectors
// LowLevelMultipleAddMultiply over an array of vectors (each its own allocation, as DiagHam's
// RealVector* vSources): up to 16 of them are interleaved into one [row][k] block, multiplied by the
// block kernels (one index lookup per element for the whole block) and written back.
struct VectorPointers {
  const double* x[16];
  double* y[16];
};

template <int K>
__global__ void interleave_kernel(VectorPointers v, double* __restrict__ X, int64_t dim) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < dim; r += stride) {
    double t[K];
#pragma unroll
    for (int q = 0; q < K; ++q) t[q] = v.x[q][r];
#pragma unroll
    for (int q = 0; q < K; ++q) X[r * K + q] = t[q];
  }
}

template <int K>
__global__ void deinterleave_kernel(VectorPointers v, const double* __restrict__ Y, int64_t first, int64_t count,
                                    int add) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    const int64_t r = first + i;
#pragma unroll
    for (int q = 0; q < K; ++q) v.y[q][r] = (add ? v.y[q][r] : 0.0) + Y[r * K + q];
  }
}

template <int K>
static int hv_group(dh_hamiltonian* h, const double* const* x, double* const* y, int add, int64_t first,
                    int64_t count) {
  dh_context* c = h->ctx;
  const int64_t dim = h->basis->dim;
  const size_t need = sizeof(double) * (size_t)dim * K * 2;
  if (need > c->packBytes) {
    if (c->d_pack) {
      DH_HIP(c, hipStreamSynchronize(c->stream));
      hipFree(c->d_pack);
      c->d_pack = nullptr;
      c->packBytes = 0;
    }
    int rc = dh_alloc(c, (void**)&c->d_pack, need);
    if (rc) return rc;
    c->packBytes = need;
  }
  double* X = c->d_pack;
  double* Y = c->d_pack + (size_t)dim * K;
  VectorPointers v{};
  for (int q = 0; q < K; ++q) {
    v.x[q] = x[q];
    v.y[q] = y[q];
  }
  const unsigned grid = (unsigned)std::min<int64_t>((dim + 255) / 256, (int64_t)c->nbrCU * 8);
  hipLaunchKernelGGL(interleave_kernel<K>, dim3(std::max(grid, 1u)), dim3(256), 0, c->stream, v, X, dim);
  DH_HIP(c, hipGetLastError());
  int rc = dh_hv(h, X, Y, K, 0, first, count);
  if (rc) return rc;
  const unsigned g2 = (unsigned)std::min<int64_t>((count + 255) / 256, (int64_t)c->nbrCU * 8);
  hipLaunchKernelGGL(deinterleave_kernel<K>, dim3(std::max(g2, 1u)), dim3(256), 0, c->stream, v, Y, first, count,
                     add);
  DH_HIP(c, hipGetLastError());
  return DH_OK;
}

extern "C" int dh_hv_vectors(dh_hamiltonian* h, const double* const* x, double* const* y, int nbrVectors, int add,
                             int64_t first, int64_t count) {
  if (!h || !x || !y || nbrVectors < 0) return DH_E_ARG;
  dh_context* c = h->ctx;
  for (int q = 0; q < nbrVectors; ++q)
    if (!x[q] || !y[q]) return dh_fail(c, DH_E_ARG, "null vector in dh_hv_vectors");
  const int64_t dim = h->basis->dim;
  if (count < 0) count = dim - first;
  if (first < 0 || first + count > dim) return dh_fail(c, DH_E_ARG, "row range outside the space");
  if (count == 0) return DH_OK;
  // groups of 16, 8, 4, 2 vectors share their lookups; a last single vector goes straight through
  for (int q = 0; q < nbrVectors;) {
    const int left = nbrVectors - q;
    int rc;
    if (left >= 16) {
      rc = hv_group<16>(h, x + q, y + q, add, first, count);
      q += 16;
    } else if (left >= 8) {
      rc = hv_group<8>(h, x + q, y + q, add, first, count);
      q += 8;
    } else if (left >= 4) {
      rc = hv_group<4>(h, x + q, y + q, add, first, count);
      q += 4;
    } else if (left >= 2) {
      rc = hv_group<2>(h, x + q, y + q, add, first, count);
      q += 2;
    } else {
      rc = dh_hv(h, x[q], y[q], 1, add, first, count);
      q += 1;
    }
    if (rc) return rc;
  }
  return DH_OK;
}

// ----------------------------------------------------------------------------- Hamiltonian objects
extern "C" int dh_sphere_hamiltonian_create(dh_context* c, dh_hamiltonian** out, dh_basis* b, const double* factors,
                                            const int* pairOffset, int nbrFactors) {
  if (!c || !out || !b || !factors || !pairOffset) return DH_E_ARG;
  *out = nullptr;
  if (c->closed) return dh_fail(c, DH_E_STATE, "context destroyed");
  if (b->kind != DH_BASIS_FERMION_SPHERE) return dh_fail(c, DH_E_ARG, "basis is not a fermion-on-sphere basis");
  const int nOrb = b->nbrOrbitals;
  const int nPairs = nOrb * (nOrb - 1) / 2;
  // the layout is fixed by the reference tables: check it
  int expect = 0;
  for (int m1 = 0; m1 < nOrb; ++m1)
    for (int m2 = 0; m2 < m1; ++m2) {
      int a = m1 * (m1 - 1) / 2 + m2;
      if (pairOffset[a] != expect) return dh_fail(c, DH_E_ARG, "pair_offset does not match the factor layout");
      int sum = m1 + m2;
      expect += std::min(sum, nOrb - 1) - (sum / 2 + 1) + 1;
    }
  if (pairOffset[nPairs] != expect || nbrFactors != expect)
    return dh_fail(c, DH_E_ARG, "nbr_factors does not match the factor layout");
  std::vector<double> packed;
  std::vector<int> sumBase;
  int rc = dh_pack_factors_checked(c, nOrb, factors, pairOffset, packed, sumBase);
  if (rc) return rc;
  dh_hamiltonian* h = new dh_hamiltonian();
  h->ctx = c;
  h->basis = b;
  h->kind = DH_HAM_SPHERE;
  h->nbrFactors = nbrFactors;
  h->nbrPairs = nPairs;
  h->nbrPackedFactors = (int)packed.size();
  h->nbrPairSums = (int)sumBase.size();
  if ((rc = dh_alloc(c, (void**)&h->d_factors, sizeof(double) * std::max(nbrFactors, 1))) ||
      (rc = dh_alloc(c, (void**)&h->d_pairOffset, sizeof(int) * (nPairs + 1))) ||
      (rc = dh_alloc(c, (void**)&h->d_packedFactors, sizeof(double) * std::max<size_t>(packed.size(), 1))) ||
      (rc = dh_alloc(c, (void**)&h->d_sumBase, sizeof(int) * sumBase.size()))) {
    dh_hamiltonian_destroy(h);
    return rc;
  }
  DH_HIP(c, hipMemcpyAsync(h->d_factors, factors, sizeof(double) * nbrFactors, hipMemcpyHostToDevice, c->stream));
  DH_HIP(c, hipMemcpyAsync(h->d_pairOffset, pairOffset, sizeof(int) * (nPairs + 1), hipMemcpyHostToDevice, c->stream));
  DH_HIP(c, hipMemcpyAsync(h->d_packedFactors, packed.data(), sizeof(double) * packed.size(), hipMemcpyHostToDevice,
                           c->stream));
  DH_HIP(c, hipMemcpyAsync(h->d_sumBase, sumBase.data(), sizeof(int) * sumBase.size(), hipMemcpyHostToDevice,
                           c->stream));
  DH_HIP(c, hipStreamSynchronize(c->stream));
  h->lanesPerRow = (b->dim < (int64_t)c->nbrCU * 256) ? 4 : 1;
  h->blockSize = 512;
  h->indexMode = DH_INDEX_RANK;
  h->autoLaunch = true;
  *out = h;
  return DH_OK;
}

extern "C" int dh_sphere_hamiltonian_set_channels(dh_hamiltonian* h, const double* cg, int nbrChannels,
                                                  const double* weights) {
  if (!h || !cg || !weights || nbrChannels <= 0) return DH_E_ARG;
  if (h->kind != DH_HAM_SPHERE) return dh_fail(h->ctx, DH_E_ARG, "not a sphere Hamiltonian");
  if (nbrChannels > 32) return dh_fail(h->ctx, DH_E_RANGE, "at most 32 channels");
  dh_context* c = h->ctx;
  if (h->d_cg) {
    hipFree(h->d_cg);
    h->d_cg = nullptr;
  }
  int rc = dh_alloc(c, (void**)&h->d_cg, sizeof(double) * h->nbrPairs * nbrChannels);
  if (rc) return rc;
  DH_HIP(c, hipMemcpyAsync(h->d_cg, cg, sizeof(double) * h->nbrPairs * nbrChannels, hipMemcpyHostToDevice, c->stream));
  DH_HIP(c, hipStreamSynchronize(c->stream));
  h->nbrChannels = nbrChannels;
  h->channelWeights.assign(weights, weights + nbrChannels);
  return DH_OK;
}

extern "C" int dh_spin_chain_hamiltonian_create(dh_context* c, dh_hamiltonian** out, dh_basis* b, const double* j,
                                                const double* jz, int periodic) {
  if (!c || !out || !b || !j || !jz) return DH_E_ARG;
  *out = nullptr;
  if (c->closed) return dh_fail(c, DH_E_STATE, "context destroyed");
  if (b->kind != DH_BASIS_SPIN_CHAIN) return dh_fail(c, DH_E_ARG, "basis is not a spin chain");
  dh_hamiltonian* h = new dh_hamiltonian();
  h->ctx = c;
  h->basis = b;
  h->kind = DH_HAM_SPIN;
  h->periodic = periodic ? 1 : 0;
  int nb = periodic ? b->chainLength : b->chainLength - 1;
  for (int i = 0; i < nb; ++i) {
    h->J[i] = j[i];
    h->Jz[i] = jz[i];
  }
  h->indexMode = DH_INDEX_RANK;
  *out = h;
  return DH_OK;
}

extern "C" int dh_hamiltonian_destroy(dh_hamiltonian* h) {
  if (!h) return DH_E_ARG;
  dh_context* c = h->ctx;
  hipStreamSynchronize(c->stream);
  if (h->d_factors) hipFree(h->d_factors);
  if (h->d_pairOffset) hipFree(h->d_pairOffset);
  if (h->d_packedFactors) hipFree(h->d_packedFactors);
  if (h->d_sumBase) hipFree(h->d_sumBase);
  if (h->d_cg) hipFree(h->d_cg);
  if (h->d_adjBinom) hipFree(h->d_adjBinom);
  dh_nbody_free(h);
  dh_csr_free(h);
  delete h;
  return DH_OK;
}

extern "C" int dh_hamiltonian_set_shift(dh_hamiltonian* h, double shift) {
  if (!h) return DH_E_ARG;
  h->shift = shift;
  return DH_OK;
}

extern "C" int64_t dh_hamiltonian_dimension(const dh_hamiltonian* h) { return h ? h->basis->dim : -1; }

extern "C" int dh_hamiltonian_set_launch(dh_hamiltonian* h, int lpr, int blockSize, int mode, int lds) {
  if (!h) return DH_E_ARG;
  if (lpr != 1 && lpr != 2 && lpr != 4 && lpr != 8) return dh_fail(h->ctx, DH_E_ARG, "lanes_per_row in {1,2,4,8}");
  if (h->kind == DH_HAM_SPIN && lpr != 1) return dh_fail(h->ctx, DH_E_ARG, "spin kernels use one row per lane");
  if (blockSize < 64 || blockSize > 1024 || (blockSize % 64)) return dh_fail(h->ctx, DH_E_ARG, "block_size");
  if (mode != DH_INDEX_BSEARCH && mode != DH_INDEX_RANK && mode != DH_INDEX_RANK_X4 && mode != DH_INDEX_RANK_FREE &&
      mode != DH_INDEX_RANK_RANGE)
    return dh_fail(h->ctx, DH_E_ARG, "index_mode");
  h->lanesPerRow = lpr;
  h->blockSize = blockSize;
  h->indexMode = mode;
  h->ldsTables = lds;
  h->autoLaunch = false;
  return DH_OK;
}
