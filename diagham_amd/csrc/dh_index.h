// Device index lookups (FindStateIndex) shared by the product kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

// FermionOnSphere::FindStateIndex (FQHE/src/HilbertSpace/FermionOnSphere.cc:1431) as a
// lookup-table-seeded binary search: the LUT over the top bits of the state (staged in
// LDS by the caller) bounds the range [lut[key+1], lut[key]) of states sharing those bits,
// which is then bisected in the descending basis.
__device__ __forceinline__ int64_t dh_find_bsearch(uint64_t target, const int* __restrict__ lut, int shift,
                                                   const uint64_t* __restrict__ states) {
  uint32_t key = (uint32_t)(target >> shift);
  int64_t lo = lut[key + 1];
  int64_t hi = lut[key];
  while (hi - lo > 1) {
    int64_t mid = (lo + hi) >> 1;
    if (states[mid] >= target)
      lo = mid;
    else
      hi = mid;
  }
  return lo;  // in [lut[key+1], lut[key]) for every target of the sector
}

// Combinatorial rank of a fermion state in the descending basis of its sector (the low table in 32
// bits, or its 16-bit copy).
template <typename L>
__device__ __forceinline__ int64_t dh_find_rank(uint64_t target, const L* __restrict__ lo,
                                                const uint32_t* __restrict__ hi, int bits, int64_t dim) {
  uint32_t a = lo[target & ((1ull << bits) - 1)];
  uint32_t b = hi[target >> bits];
  // clamped so that a target outside the sector can never address outside the vectors
  int64_t r = dim - 1 - (int64_t)a - (int64_t)b;
  return r < 0 ? 0 : (r >= dim ? dim - 1 : r);
}

// Spin1_2Chain order: key = bit-reversed L-bit word (site 0 most significant)
__device__ __forceinline__ uint32_t dh_spin_key(uint32_t s, int L) { return __builtin_bitreverse32(s) >> (32 - L); }

// Spin1_2Chain::FindStateIndex (Spin/src/HilbertSpace/Spin1_2Chain.cc:1002) as a
// LUT-seeded binary search over the bit-reversed keys (s holds the L site bits only).
__device__ __forceinline__ int64_t dh_spin_find_bsearch(uint32_t s, int L, const int* __restrict__ lut, int shift,
                                                        const uint32_t* __restrict__ words) {
  uint32_t key = dh_spin_key(s, L);
  uint32_t top = key >> shift;
  int64_t lo = lut[top + 1];
  int64_t hi = lut[top];
  const uint32_t mask = (L == 32) ? 0xffffffffu : ((1u << L) - 1u);
  while (hi - lo > 1) {
    int64_t mid = (lo + hi) >> 1;
    if (dh_spin_key(words[mid] & mask, L) >= key)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Rank of a spin state: dim - 1 - sum_{up sites i} C(L-1-i, K_i), K_i = ups at sites >= i.
__device__ __forceinline__ int64_t dh_spin_find_rank(uint32_t s, int L, const uint32_t* __restrict__ binom,
                                                     int64_t dim) {
  int64_t less = 0;
  for (uint32_t r = s; r; r &= r - 1) {
    int i = __builtin_ctz(r);
    int K = __popc(s >> i);
    less += binom[(L - 1 - i) * 33 + K];
  }
  int64_t r = dim - 1 - less;
  return r < 0 ? 0 : (r >= dim ? dim - 1 : r);
}
