// Per-row accumulation of the fermions-on-the-sphere product, shared by the single-sector
// and the multi-sector kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dh_index.h"
#include "dh_internal.h"

// Bit helpers over the state word: 64-bit words in general, 32-bit words when the orbitals fit
// (half the instructions for every shift, mask, ctz and popcount of the row routine).
__device__ __forceinline__ int dh_ctz(uint64_t v) { return __builtin_ctzll(v); }
__device__ __forceinline__ int dh_ctz(uint32_t v) { return __builtin_ctz(v); }
__device__ __forceinline__ int dh_popc(uint64_t v) { return __popcll(v); }
__device__ __forceinline__ int dh_popc(uint32_t v) { return __popc(v); }

// dh_packed_index with a 24-bit multiply (positions stay below 2^12): a full-rate multiply-add
// instead of the quarter-rate 32-bit multiply
__device__ __forceinline__ int dh_packed_index_dev(int i, int j) {
  const unsigned mx = (unsigned)max(i, j), mn = (unsigned)min(i, j);
  return (int)((__umul24(mx, mx + 1u) >> 1) + mn);
}

// c, negated when the parity p is odd: the sign bit flipped (the same double as -c)
__device__ __forceinline__ double dh_negate_if(double c, int p) {
  return __hiloint2double(__double2hiint(c) ^ (int)((unsigned)p << 31), __double2loint(c));
}

template <int K, typename I>
__device__ __forceinline__ void dh_gather_fma(double* acc, const double* __restrict__ x, I idx, double coef) {
  const double* xr = x + (size_t)idx * K;
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

// Two-body terms of row state s: for every occupied pair n1 > n2 (annihilated as a_n1 a_n2,
// sign = parity of the occupied orbitals at and above each), every created pair m3 > m4 with
// m3 + m4 = n1 + n2 (created a+_m3 a+_m4 on the intermediate state) whose orbitals are free.
// Lane `lane` of the row's LPR lanes takes m3 = m3lo + lane, m3lo + lane + LPR, ...
// Factors come packed symmetric per pair sum (sF, sBase: see dh_pack_sphere_factors).
// Target indices come from `find(target)`; x is indexed relative to `xbase`.  Each term is one
// gather.  For K <= 2 the multiply-add of a term is deferred to the next term, after that term's
// index loads are issued: a term's gather and the next term's index loads are then in flight
// together, instead of one after the other (the terms are still added in their order, so the sum
// is the same to the bit).  Wider K keep one gather in flight per term: their registers would
// lower the occupancy that hides the latency there.
template <int K, int LPR, typename U, typename Find>
__device__ __forceinline__ void dh_sphere_row(U s, int nbrOrbitals, const double* __restrict__ sF,
                                              const int* __restrict__ sBase, int lane, Find find,
                                              const double* __restrict__ xbase, double* acc) {
  constexpr bool PIPE = K <= 2;
  double pend[K];
  double pendCoef = 0.0;
#pragma unroll
  for (int v = 0; v < K; ++v) pend[v] = 0.0;
  for (U r1 = s; r1; r1 &= r1 - 1) {
    const int n1 = dh_ctz(r1);
    const int pc1 = dh_popc(s >> n1);
    for (U r2 = s & ((U(1) << n1) - U(1)); r2; r2 &= r2 - 1) {
      const int n2 = dh_ctz(r2);
      const int sgnA = pc1 + dh_popc(s >> n2);
      const U t = s ^ (U(1) << n1) ^ (U(1) << n2);
      const int sum = n1 + n2;
      const int m3lo = (sum >> 1) + 1;
      const int m3hi = min(sum, nbrOrbitals - 1);
      const double* fS = sF + sBase[sum];
      const int i = n1 - m3lo;
      for (int m3 = m3lo + lane; m3 <= m3hi; m3 += LPR) {
        const int m4 = sum - m3;
        const U cm = (U(1) << m3) | (U(1) << m4);
        if (t & cm) continue;
        const int sg = sgnA + dh_popc(t >> m4) + dh_popc(t >> m3);
        const auto idx = find(t | cm);
        double coef = fS[dh_packed_index_dev(i, m3 - m3lo)];
        if (sg & 1) coef = -coef;
        if constexpr (PIPE) {
          // the previous term first, then this term's gather into the same registers (no copy,
          // so nothing waits for the gather before the next term's index loads)
          const double* xr = xbase + (size_t)idx * K;
#pragma unroll
          for (int v = 0; v < K; ++v) acc[v] = fma(pendCoef, pend[v], acc[v]);
#pragma unroll
          for (int v = 0; v < K; ++v) pend[v] = xr[v];
          pendCoef = coef;
        } else {
          dh_gather_fma<K>(acc, xbase, idx, coef);
        }
      }
    }
  }
  if constexpr (PIPE) {
#pragma unroll
    for (int v = 0; v < K; ++v) acc[v] = fma(pendCoef, pend[v], acc[v]);
  }
}

// The same terms with the created-pair loop unrolled four-wide: a lane computes four targets
// (a slot past the last pair or on an occupied orbital looks up the row's own state with a zero
// coefficient), issues their index lookups together and then their four gathers, so four
// dependent chains are in flight per lane.  `s` must be a state of the space (the padding
// slots look it up).
template <int K, int LPR, typename U, typename Find>
__device__ __forceinline__ void dh_sphere_row_x4(U s, int nbrOrbitals, const double* __restrict__ sF,
                                                 const int* __restrict__ sBase, int lane, Find find,
                                                 const double* __restrict__ xbase, double* acc) {
  for (U r1 = s; r1; r1 &= r1 - 1) {
    const int n1 = dh_ctz(r1);
    const int pc1 = dh_popc(s >> n1);
    for (U r2 = s & ((U(1) << n1) - U(1)); r2; r2 &= r2 - 1) {
      const int n2 = dh_ctz(r2);
      const int sgnA = pc1 + dh_popc(s >> n2);
      const U t = s ^ (U(1) << n1) ^ (U(1) << n2);
      const int sum = n1 + n2;
      const int m3lo = (sum >> 1) + 1;
      const int m3hi = min(sum, nbrOrbitals - 1);
      const double* fS = sF + sBase[sum];
      const int i = n1 - m3lo;
      for (int m3b = m3lo + lane; m3b <= m3hi; m3b += 4 * LPR) {
        decltype(find(s)) idx[4];
        double coef[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int m3 = min(m3b + u * LPR, m3hi);  // clamped so that every shift stays in range
          const int m4 = sum - m3;
          const U cm = (U(1) << m3) | (U(1) << m4);
          const bool ok = (m3b + u * LPR <= m3hi) && !(t & cm);
          const int sg = sgnA + dh_popc(t >> m4) + dh_popc(t >> m3);
          const double c = ok ? fS[dh_packed_index_dev(i, m3 - m3lo)] : 0.0;
          coef[u] = (sg & 1) ? -c : c;
          idx[u] = find(ok ? (t | cm) : s);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) dh_gather_fma<K>(acc, xbase, idx[u], coef[u]);
      }
    }
  }
}

__device__ __forceinline__ uint64_t dh_bitrev(uint64_t v) { return __builtin_bitreverse64(v); }
__device__ __forceinline__ uint32_t dh_bitrev(uint32_t v) { return __builtin_bitreverse32(v); }

// Created pairs (m3 > m4, m3 + m4 = sum, both orbitals free in t) as a mask of their m3 bits:
// bit m of f & (f mirrored about sum / 2) says that m and sum - m are both free.  The mirror of f
// is its bit reversal within the nbrOrbitals orbitals, shifted so that bit sum - m lands on bit m.
template <typename U>
__device__ __forceinline__ U dh_free_pairs(U t, int nbrOrbitals, int sum, int m3lo, int m3hi) {
  constexpr int W = 8 * sizeof(U);
  const U full = nbrOrbitals >= W ? ~U(0) : ((U(1) << nbrOrbitals) - U(1));
  const U f = ~t & full;
  const U g = dh_bitrev(f) >> (W - nbrOrbitals);  // bit j of g = bit nbrOrbitals - 1 - j of f
  const int d = nbrOrbitals - 1 - sum;             // bit m of the mirror = bit m + d of g
  const U mirror = d >= 0 ? (g >> d) : (g << (-d));
  const U hiMask = m3hi >= W - 1 ? ~U(0) : ((U(1) << (m3hi + 1)) - U(1));
  const U loMask = (U(1) << m3lo) - U(1);
  return f & mirror & hiMask & ~loMask;
}

// The terms of dh_sphere_row in the same order, with the created pairs taken from the mask of free
// pairs of each annihilated pair (dh_free_pairs): no iteration is spent on a created pair with an
// occupied orbital.  With LPR lanes on a row, lane l takes the free pairs l, l + LPR, ... of each
// annihilated pair.  For K <= 2 a term's multiply-add waits for the next term's index loads, as in
// dh_sphere_row, so the terms are added in the same order and with one lane per row the sum is
// bitwise dh_sphere_row's.
template <int K, int LPR, typename U, typename Find>
__device__ __forceinline__ void dh_sphere_row_free(U s, int nbrOrbitals, const double* __restrict__ sF,
                                                   const int* __restrict__ sBase, int lane, Find find,
                                                   const double* __restrict__ xbase, double* acc) {
  constexpr bool PIPE = K <= 2;
  double pend[K];
  double pendCoef = 0.0;
#pragma unroll
  for (int v = 0; v < K; ++v) pend[v] = 0.0;
  // one created pair m3 > m4 = sum - m3 of the intermediate state t, with factor c
  auto term = [&](U t, int sum, int sgnA, int m3, double c) {
    const U b3 = U(1) << m3, b4 = U(1) << (sum - m3);
    // the occupied orbitals of t at or above m4 and at or above m3 differ by those in [m4, m3)
    const double coef = dh_negate_if(c, sgnA + dh_popc(t & (b3 - b4)));
    const auto idx = find(t | b3 | b4);
    if constexpr (PIPE) {
      const double* xr = xbase + (size_t)idx * K;
#pragma unroll
      for (int v = 0; v < K; ++v) acc[v] = fma(pendCoef, pend[v], acc[v]);
#pragma unroll
      for (int v = 0; v < K; ++v) pend[v] = xr[v];
      pendCoef = coef;
    } else {
      dh_gather_fma<K>(acc, xbase, idx, coef);
    }
  };
  for (U r1 = s; r1; r1 &= r1 - 1) {
    const int n1 = dh_ctz(r1);
    const int pc1 = dh_popc(s >> n1);
    for (U r2 = s & ((U(1) << n1) - U(1)); r2; r2 &= r2 - 1) {
      const int n2 = dh_ctz(r2);
      const int sgnA = pc1 + dh_popc(s >> n2);
      const U t = s ^ (U(1) << n1) ^ (U(1) << n2);
      const int sum = n1 + n2;
      const int m3lo = (sum >> 1) + 1;
      const double* fS = sF + sBase[sum];
      const int i = n1 - m3lo;
      U pm = dh_free_pairs(t, nbrOrbitals, sum, m3lo, min(sum, nbrOrbitals - 1));
      if constexpr (LPR > 1)
        for (int q = 0; q < lane; ++q) pm &= pm - 1;
      while (pm) {
        const int m3 = dh_ctz(pm);
#pragma unroll
        for (int q = 0; q < LPR; ++q) pm &= pm - 1;  // this term, then the other lanes' next ones
        term(t, sum, sgnA, m3, fS[dh_packed_index_dev(i, m3 - m3lo)]);
      }
    }
  }
  if constexpr (PIPE) {
#pragma unroll
    for (int v = 0; v < K; ++v) acc[v] = fma(pendCoef, pend[v], acc[v]);
  }
}
