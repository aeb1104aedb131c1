// Index lookup, matrix-element enumeration and the adjoint (bilinear) kernel.
#include <algorithm>
#include <cstring>

#include "dh_index.h"
#include "dh_internal.h"
#include "dh_sphere_row.h"

// ----------------------------------------------------------------------------- FindStateIndex
__global__ void fs_find_kernel(const uint64_t* __restrict__ in, int64_t n, int64_t* __restrict__ out,
                               const uint64_t* __restrict__ states, int64_t dim, int nbrOrbitals, int N, int S,
                               int mode, const int* __restrict__ lut, int lutShift, const uint32_t* __restrict__ lo,
                               const uint32_t* __restrict__ hi, int rankBits) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t s = in[i];
  int sum = 0;
  for (uint64_t r = s; r; r &= r - 1) sum += __builtin_ctzll(r);
  bool valid = (nbrOrbitals >= 64 || (s >> nbrOrbitals) == 0) && __popcll(s) == N && sum == S;
  if (!valid) {
    out[i] = dim;
    return;
  }
  out[i] = (mode == DH_INDEX_BSEARCH) ? dh_find_bsearch(s, lut, lutShift, states) : dh_find_rank(s, lo, hi, rankBits, dim);
}

__global__ void spin_find_kernel(const uint64_t* __restrict__ in, int64_t n, int64_t* __restrict__ out,
                                 const uint32_t* __restrict__ words, int64_t dim, int L, int nbrUp, int mode,
                                 const int* __restrict__ lut, int lutShift, const uint32_t* __restrict__ binom) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t w = in[i];
  uint32_t mask = (L == 32) ? 0xffffffffu : ((1u << L) - 1u);
  uint32_t s = (uint32_t)w & mask;
  // reference words keep the bits L..31 at 1 and nothing above bit 31
  bool valid = (w >> 32) == 0 && (((uint32_t)w | mask) == 0xffffffffu) && __popc(s) == nbrUp;
  if (!valid) {
    out[i] = dim;
    return;
  }
  out[i] = (mode == DH_INDEX_BSEARCH) ? dh_spin_find_bsearch(s, L, lut, lutShift, words)
                                     : dh_spin_find_rank(s, L, binom, dim);
}

extern "C" int dh_basis_find_state_index(dh_basis* b, const uint64_t* states, int64_t n, int mode, int64_t* out) {
  if (!b || (n && (!states || !out)) || n < 0) return DH_E_ARG;
  if (mode != DH_INDEX_BSEARCH && mode != DH_INDEX_RANK) return dh_fail(b->ctx, DH_E_ARG, "mode");
  if (n == 0) return DH_OK;
  dh_context* c = b->ctx;
  uint64_t* d_in;
  int64_t* d_out;
  int rc;
  if ((rc = dh_alloc(c, (void**)&d_in, sizeof(uint64_t) * n))) return rc;
  if ((rc = dh_alloc(c, (void**)&d_out, sizeof(int64_t) * n))) {
    hipFree(d_in);
    return rc;
  }
  hipMemcpyAsync(d_in, states, sizeof(uint64_t) * n, hipMemcpyHostToDevice, c->stream);
  unsigned blocks = (unsigned)((n + 255) / 256);
  if (b->kind == DH_BASIS_FERMION_SPHERE)
    hipLaunchKernelGGL(fs_find_kernel, dim3(blocks), dim3(256), 0, c->stream, d_in, n, d_out, b->d_states, b->dim,
                       b->nbrOrbitals, b->nbrFermions, b->orbitalSum, mode, b->d_lut, b->lutShift, b->d_rankLo,
                       b->d_rankHi, b->rankBits);
  else
    hipLaunchKernelGGL(spin_find_kernel, dim3(blocks), dim3(256), 0, c->stream, d_in, n, d_out, b->d_words, b->dim,
                       b->chainLength, b->nbrUp, mode, b->d_lut, b->lutShift, b->d_binom);
  hipMemcpyAsync(out, d_out, sizeof(int64_t) * n, hipMemcpyDeviceToHost, c->stream);
  hipError_t e = hipStreamSynchronize(c->stream);
  hipFree(d_in);
  hipFree(d_out);
  if (e != hipSuccess) return dh_fail(c, DH_E_HIP, hipGetErrorString(e));
  return DH_OK;
}

// ----------------------------------------------------------------------------- element enumeration
// counts != 0, targets == 0: count the elements of each row (counts[row - rowBegin]);
// targets != 0, rowOffsets == 0: list the elements of row rowBegin (at most maxOut);
// rowOffsets != 0: store the elements of every row at rowOffsets[row] (32-bit targets in csrCol)
__global__ void fs_elements_kernel(const uint64_t* __restrict__ states, int64_t dim, int nbrOrbitals,
                                   const double* __restrict__ factors, const int* __restrict__ pairOffset,
                                   const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi, int rankBits,
                                   int64_t rowBegin, int64_t rowEnd, int64_t* __restrict__ counts,
                                   int64_t* __restrict__ targets, double* __restrict__ coefs, int64_t maxOut,
                                   const int64_t* __restrict__ rowOffsets = nullptr,
                                   int32_t* __restrict__ csrCol = nullptr) {
  int64_t row = rowBegin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= rowEnd) return;
  uint64_t s = states[row];
  int64_t n = 0;
  if (rowOffsets) {
    const int64_t base = rowOffsets[row];
    for (uint64_t r1 = s; r1; r1 &= r1 - 1) {
      int n1 = __builtin_ctzll(r1);
      int pc1 = __popcll(s >> n1);
      for (uint64_t r2 = s & ((1ull << n1) - 1ull); r2; r2 &= r2 - 1) {
        int n2 = __builtin_ctzll(r2);
        int sgnA = pc1 + __popcll(s >> n2);
        uint64_t t = s ^ (1ull << n1) ^ (1ull << n2);
        int sum = n1 + n2;
        int m3lo = (sum >> 1) + 1;
        int m3hi = min(sum, nbrOrbitals - 1);
        int fbase = pairOffset[(n1 * (n1 - 1)) / 2 + n2] - m3lo;
        for (int m3 = m3lo; m3 <= m3hi; ++m3) {
          int m4 = sum - m3;
          uint64_t cm = (1ull << m3) | (1ull << m4);
          if (t & cm) continue;
          int sg = sgnA + __popcll(t >> m4) + __popcll(t >> m3);
          double coef = factors[fbase + m3];
          csrCol[base + n] = (int32_t)dh_find_rank(t | cm, lo, hi, rankBits, dim);
          coefs[base + n] = (sg & 1) ? -coef : coef;
          ++n;
        }
      }
    }
    return;
  }
  for (uint64_t r1 = s; r1; r1 &= r1 - 1) {
    int n1 = __builtin_ctzll(r1);
    int pc1 = __popcll(s >> n1);
    for (uint64_t r2 = s & ((1ull << n1) - 1ull); r2; r2 &= r2 - 1) {
      int n2 = __builtin_ctzll(r2);
      int sgnA = pc1 + __popcll(s >> n2);
      uint64_t t = s ^ (1ull << n1) ^ (1ull << n2);
      int sum = n1 + n2;
      int m3lo = (sum >> 1) + 1;
      int m3hi = min(sum, nbrOrbitals - 1);
      int fbase = pairOffset[(n1 * (n1 - 1)) / 2 + n2] - m3lo;
      for (int m3 = m3lo; m3 <= m3hi; ++m3) {
        int m4 = sum - m3;
        uint64_t cm = (1ull << m3) | (1ull << m4);
        if (t & cm) continue;
        if (targets) {
          int sg = sgnA + __popcll(t >> m4) + __popcll(t >> m3);
          double coef = factors[fbase + m3];
          if (n < maxOut) {
            targets[n] = dh_find_rank(t | cm, lo, hi, rankBits, dim);
            coefs[n] = (sg & 1) ? -coef : coef;
          }
        }
        ++n;
      }
    }
  }
  if (counts) counts[row - rowBegin] = n;
  if (targets && row == rowBegin) counts[0] = n;
}

__global__ void spin_elements_kernel(const uint32_t* __restrict__ words, int64_t dim, int L, int periodic,
                                     const uint32_t* __restrict__ binom, const double* __restrict__ halfJ,
                                     int64_t rowBegin, int64_t rowEnd, int64_t* __restrict__ counts,
                                     int64_t* __restrict__ targets, double* __restrict__ coefs, int64_t maxOut,
                                     double diagValue) {
  int64_t row = rowBegin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= rowEnd) return;
  uint32_t mask = (L == 32) ? 0xffffffffu : ((1u << L) - 1u);
  uint32_t s = words[row] & mask;
  int64_t n = 0;
  if (targets && n < maxOut) {
    targets[n] = row;
    coefs[n] = diagValue;
  }
  ++n;  // the diagonal element
  for (int i = 0; i < L - 1; ++i) {
    uint32_t pr = (s >> i) & 3u;
    if (pr == 1u || pr == 2u) {
      if (targets && n < maxOut) {
        targets[n] = dh_spin_find_rank(s ^ (3u << i), L, binom, dim);
        coefs[n] = halfJ[i];
      }
      ++n;
    }
  }
  if (periodic && (((s >> (L - 1)) ^ s) & 1u)) {
    if (targets && n < maxOut) {
      targets[n] = dh_spin_find_rank(s ^ (1u | (1u << (L - 1))), L, binom, dim);
      coefs[n] = halfJ[L - 1];
    }
    ++n;
  }
  if (counts) counts[row - rowBegin] = n;
  if (targets && row == rowBegin) counts[0] = n;
}

__global__ void sum_i64_kernel(const int64_t* __restrict__ v, int64_t n, unsigned long long* __restrict__ out) {
  __shared__ unsigned long long red[1024 / 64];
  unsigned long long s = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s += (unsigned long long)v[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (unsigned w = 0; w < blockDim.x / 64; ++w) t += red[w];
    atomicAdd(out, t);
  }
}

static double spin_diag_value(const dh_hamiltonian* h, uint32_t s) {
  int L = h->basis->chainLength;
  double d = h->shift;
  for (int b = 0; b < L - 1; ++b) d += h->Jz[b] * ((((s >> b) ^ (s >> (b + 1))) & 1u) ? -0.25 : 0.25);
  if (h->periodic) d += h->Jz[L - 1] * ((((s >> (L - 1)) ^ s) & 1u) ? -0.25 : 0.25);
  return d;
}

extern "C" int dh_hamiltonian_count_elements(dh_hamiltonian* h, int64_t* nnz) {
  if (!h || !nnz) return DH_E_ARG;
  if (h->kind == DH_HAM_SPHERE_NBODY) return dh_nbody_count_elements(h, nnz);
  dh_context* c = h->ctx;
  dh_basis* b = h->basis;
  const int64_t chunk = 1 << 24;
  int64_t* d_counts;
  unsigned long long* d_total;
  int rc;
  if ((rc = dh_alloc(c, (void**)&d_counts, sizeof(int64_t) * std::min<int64_t>(chunk, b->dim)))) return rc;
  if ((rc = dh_alloc(c, (void**)&d_total, sizeof(unsigned long long)))) {
    hipFree(d_counts);
    return rc;
  }
  hipMemsetAsync(d_total, 0, sizeof(unsigned long long), c->stream);
  double* d_halfJ = nullptr;
  if (h->kind == DH_HAM_SPIN) {
    double hj[32];
    for (int i = 0; i < 32; ++i) hj[i] = 0.5 * h->J[i];
    dh_alloc(c, (void**)&d_halfJ, sizeof(hj));
    hipMemcpyAsync(d_halfJ, hj, sizeof(hj), hipMemcpyHostToDevice, c->stream);
  }
  for (int64_t r0 = 0; r0 < b->dim; r0 += chunk) {
    int64_t r1 = std::min(b->dim, r0 + chunk);
    unsigned blocks = (unsigned)((r1 - r0 + 255) / 256);
    if (h->kind == DH_HAM_SPHERE)
      hipLaunchKernelGGL(fs_elements_kernel, dim3(blocks), dim3(256), 0, c->stream, b->d_states, b->dim,
                         b->nbrOrbitals, h->d_factors, h->d_pairOffset, b->d_rankLo, b->d_rankHi, b->rankBits, r0, r1,
                         d_counts, (int64_t*)nullptr, (double*)nullptr, (int64_t)0);
    else
      hipLaunchKernelGGL(spin_elements_kernel, dim3(blocks), dim3(256), 0, c->stream, b->d_words, b->dim,
                         b->chainLength, h->periodic, b->d_binom, d_halfJ, r0, r1, d_counts, (int64_t*)nullptr,
                         (double*)nullptr, (int64_t)0, 0.0);
    hipLaunchKernelGGL(sum_i64_kernel, dim3(256), dim3(256), 0, c->stream, d_counts, r1 - r0, d_total);
  }
  unsigned long long total = 0;
  hipMemcpyAsync(&total, d_total, sizeof(total), hipMemcpyDeviceToHost, c->stream);
  hipError_t e = hipStreamSynchronize(c->stream);
  hipFree(d_counts);
  hipFree(d_total);
  if (d_halfJ) hipFree(d_halfJ);
  if (e != hipSuccess) return dh_fail(c, DH_E_HIP, hipGetErrorString(e));
  *nnz = (int64_t)total;
  return DH_OK;
}

extern "C" int dh_sector_set_count_elements(dh_sector_set* S, int64_t* counts) {
  if (!S || !counts) return DH_E_ARG;
  dh_context* c = S->ctx;
  if (S->kind == DH_SECTORS_MOMENTUM) return dh_lattice_count_elements(S, counts);
  if (S->kind != DH_SECTORS_SPHERE)
    return dh_fail(c, DH_E_ARG, "element counts are implemented for sphere and momentum sector sets");
  if (!S->d_factors) return dh_fail(c, DH_E_STATE, "factors not set");
  const int64_t chunk = 1 << 24;
  int64_t maxDim = 0;
  for (int s = 0; s < S->nbrSectors; ++s) maxDim = std::max(maxDim, S->offsets[s + 1] - S->offsets[s]);
  int64_t* d_counts = nullptr;
  unsigned long long* d_total = nullptr;
  int rc;
  if ((rc = dh_alloc(c, (void**)&d_counts, sizeof(int64_t) * std::max<int64_t>(1, std::min(chunk, maxDim)))))
    return rc;
  if ((rc = dh_alloc(c, (void**)&d_total, sizeof(unsigned long long) * S->nbrSectors))) {
    hipFree(d_counts);
    return rc;
  }
  hipMemsetAsync(d_total, 0, sizeof(unsigned long long) * S->nbrSectors, c->stream);
  for (int s = 0; s < S->nbrSectors; ++s) {
    const int64_t off = S->offsets[s], dim = S->offsets[s + 1] - off;
    for (int64_t r0 = 0; r0 < dim; r0 += chunk) {
      const int64_t r1 = std::min(dim, r0 + chunk);
      const unsigned blocks = (unsigned)((r1 - r0 + 255) / 256);
      // counting only: the target lookup tables are not read
      hipLaunchKernelGGL(fs_elements_kernel, dim3(blocks), dim3(256), 0, c->stream, S->d_states + off, dim,
                         S->nbrOrbitals, S->d_factors, S->d_pairOffset, (const uint32_t*)nullptr,
                         (const uint32_t*)nullptr, 0, r0, r1, d_counts, (int64_t*)nullptr, (double*)nullptr,
                         (int64_t)0);
      hipLaunchKernelGGL(sum_i64_kernel, dim3(256), dim3(256), 0, c->stream, d_counts, r1 - r0, d_total + s);
    }
  }
  std::vector<unsigned long long> total(S->nbrSectors);
  hipMemcpyAsync(total.data(), d_total, sizeof(unsigned long long) * S->nbrSectors, hipMemcpyDeviceToHost,
                 c->stream);
  hipError_t e = hipStreamSynchronize(c->stream);
  hipFree(d_counts);
  hipFree(d_total);
  if (e != hipSuccess) return dh_fail(c, DH_E_HIP, hipGetErrorString(e));
  for (int s = 0; s < S->nbrSectors; ++s) counts[s] = (int64_t)total[s];
  return DH_OK;
}

extern "C" int dh_hamiltonian_row_elements(dh_hamiltonian* h, int64_t row, int64_t* targets, double* coefs,
                                           int64_t maxOut, int64_t* n) {
  if (!h || !n || (maxOut > 0 && (!targets || !coefs)) || maxOut < 0) return DH_E_ARG;
  dh_context* c = h->ctx;
  dh_basis* b = h->basis;
  if (row < 0 || row >= b->dim) return dh_fail(c, DH_E_ARG, "row outside the space");
  if (h->kind == DH_HAM_SPHERE_NBODY)
    return dh_fail(c, DH_E_ARG, "row listing is implemented for two-body and spin-chain Hamiltonians");
  int64_t cap = std::max<int64_t>(maxOut, 1);
  int64_t *d_t, *d_n;
  double* d_c;
  double* d_halfJ = nullptr;
  dh_alloc(c, (void**)&d_t, sizeof(int64_t) * cap);
  dh_alloc(c, (void**)&d_c, sizeof(double) * cap);
  dh_alloc(c, (void**)&d_n, sizeof(int64_t));
  if (h->kind == DH_HAM_SPHERE) {
    hipLaunchKernelGGL(fs_elements_kernel, dim3(1), dim3(64), 0, c->stream, b->d_states, b->dim, b->nbrOrbitals,
                       h->d_factors, h->d_pairOffset, b->d_rankLo, b->d_rankHi, b->rankBits, row, row + 1, d_n, d_t,
                       d_c, maxOut);
  } else {
    double hj[32];
    for (int i = 0; i < 32; ++i) hj[i] = 0.5 * h->J[i];
    dh_alloc(c, (void**)&d_halfJ, sizeof(hj));
    hipMemcpyAsync(d_halfJ, hj, sizeof(hj), hipMemcpyHostToDevice, c->stream);
    uint64_t w = 0;
    hipMemcpyAsync(&w, b->d_words + row, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream);
    hipStreamSynchronize(c->stream);
    uint32_t mask = (b->chainLength == 32) ? 0xffffffffu : ((1u << b->chainLength) - 1u);
    double dv = spin_diag_value(h, (uint32_t)w & mask);
    hipLaunchKernelGGL(spin_elements_kernel, dim3(1), dim3(64), 0, c->stream, b->d_words, b->dim, b->chainLength,
                       h->periodic, b->d_binom, d_halfJ, row, row + 1, d_n, d_t, d_c, maxOut, dv);
  }
  int64_t cnt = 0;
  hipMemcpyAsync(&cnt, d_n, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream);
  hipStreamSynchronize(c->stream);
  int64_t m = std::min(cnt, maxOut);
  if (m > 0) {
    hipMemcpyAsync(targets, d_t, sizeof(int64_t) * m, hipMemcpyDeviceToHost, c->stream);
    hipMemcpyAsync(coefs, d_c, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream);
  }
  hipError_t e = hipStreamSynchronize(c->stream);
  hipFree(d_t);
  hipFree(d_c);
  hipFree(d_n);
  if (d_halfJ) hipFree(d_halfJ);
  if (e != hipSuccess) return dh_fail(c, DH_E_HIP, hipGetErrorString(e));
  *n = cnt;
  return DH_OK;
}

// ----------------------------------------------------------------------------- adjoint
// With the interaction split by channel, F_c(a -> b) = -4 u_c[a] u_c[b] (u_c[p] the Clebsch-Gordan
// coefficient of pair p in channel c) is of rank one in the pairs, so the bilinear factorises over
// the (N-2)-particle states t that a pair annihilation leaves:
//   g[c] = <psi|H_c|phi> = 4 sum_t A_c(psi)[t] A_c(phi)[t],
//   A_c(v)[t] = sum_{pairs p free in t, sum(p) = sum(s) - sum(t)} u_c[p] (-1)^(popc(t >> p_lo) + popc(t >> p_hi)) v[t | p]
// (the annihilation sign of a pair is its creation sign plus one, and the two cancel).  Each (row,
// annihilated pair, created pair) term of the product becomes one (t, pair) term here, so a
// bilinear costs C(N, 2) gathers per row instead of one per matrix element.  One lane per t, the
// t unranked from its index in the combinatorial number system over all (N-2)-subsets of the
// orbitals; the index loads, then the gathers, of four pairs are in flight together.  SAME: psi and
// phi are the same vector (an expectation value), one sum per channel.  Channels [c0, c0 + NC).
template <int NC, bool SAME>
__global__ void __launch_bounds__(256) sphere_adjoint_kernel(int nbrOrbitals, int holes, int64_t nbrSubsets,
                                                             int orbitalSum, const int64_t* __restrict__ binom,
                                                             int nbrChannels, int c0, const double* __restrict__ cg,
                                                             int nbrPairs, const uint32_t* __restrict__ lo,
                                                             const uint32_t* __restrict__ hi, int rankBits,
                                                             int64_t dim, const double* __restrict__ psi,
                                                             const double* __restrict__ phi,
                                                             double* __restrict__ partial) {
  extern __shared__ double smem[];
  double* sU = smem;                                                 // [nbrPairs][NC]
  int64_t* sB = reinterpret_cast<int64_t*>(smem + nbrPairs * NC);    // [nbrOrbitals + 1][holes + 1]
  __shared__ double red[256 / 64][NC];
  const int nc = min(NC, nbrChannels - c0);
  for (int i = threadIdx.x; i < nbrPairs * NC; i += blockDim.x) {
    const int p = i / NC, c = i - p * NC;
    sU[i] = c < nc ? cg[(int64_t)p * nbrChannels + c0 + c] : 0.0;
  }
  for (int i = threadIdx.x; i < (nbrOrbitals + 1) * (holes + 1); i += blockDim.x) sB[i] = binom[i];
  __syncthreads();
  double g[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) g[c] = 0.0;
  const int maxSum = 2 * nbrOrbitals - 3;
  for (int64_t it = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; it < nbrSubsets;
       it += (int64_t)gridDim.x * blockDim.x) {
    // the it-th (N-2)-subset: it = sum_k C(c_k, k) over its orbitals c_1 < ... < c_m
    uint64_t t = 0;
    int64_t r = it;
    int m = holes, sumT = 0;
    for (int j = nbrOrbitals - 1; j >= 0 && m > 0; --j) {
      const int64_t b = sB[j * (holes + 1) + m];
      if (r >= b) {
        t |= 1ull << j;
        sumT += j;
        r -= b;
        --m;
      }
    }
    const int S = orbitalSum - sumT;  // the pair sum that completes t to the sector
    if (S < 1 || S > maxSum) continue;
    uint64_t pm = dh_free_pairs(t, nbrOrbitals, S, (S >> 1) + 1, min(S, nbrOrbitals - 1));
    if (!pm) continue;
    double a[NC], bsum[SAME ? 1 : NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) a[c] = 0.0;
    if constexpr (!SAME) {
#pragma unroll
      for (int c = 0; c < NC; ++c) bsum[c] = 0.0;
    }
    while (pm) {
      int64_t idx[4];
      int pb[4];
      bool neg[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        idx[q] = 0;  // a slot past the last pair reads row 0 with no weight
        pb[q] = -1;
        neg[q] = false;
        if (pm) {
          const int m3 = __builtin_ctzll(pm);
          pm &= pm - 1;
          const int m4 = S - m3;
          idx[q] = dh_find_rank(t | (1ull << m3) | (1ull << m4), lo, hi, rankBits, dim);
          pb[q] = (m3 * (m3 - 1)) / 2 + m4;
          neg[q] = (__popcll(t >> m4) + __popcll(t >> m3)) & 1;
        }
      }
      double x[4], z[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        x[q] = psi[idx[q]];
        if constexpr (!SAME) z[q] = phi[idx[q]];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (pb[q] < 0) continue;
        const double* u = sU + (int64_t)pb[q] * NC;
        const double xs = neg[q] ? -x[q] : x[q];
#pragma unroll
        for (int c = 0; c < NC; ++c) a[c] += u[c] * xs;
        if constexpr (!SAME) {
          const double zs = neg[q] ? -z[q] : z[q];
#pragma unroll
          for (int c = 0; c < NC; ++c) bsum[c] += u[c] * zs;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if constexpr (SAME)
        g[c] += a[c] * a[c];
      else
        g[c] += a[c] * bsum[c];
    }
  }
  const int wl = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    double v = g[c];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (wl == 0) red[wave][c] = v;
  }
  __syncthreads();
  if (threadIdx.x < NC) {
    double v = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v += red[w][threadIdx.x];
    partial[(int64_t)blockIdx.x * NC + threadIdx.x] = 4.0 * v;
  }
}

// out[c] = sum over the blocks of partial[b][c]: one 256-thread workgroup per column c, the threads
// striding over the blocks (a single lane per column would wait for a thousand dependent loads)
__global__ void __launch_bounds__(256) reduce_partials_kernel(const double* __restrict__ partial, int nbrBlocks,
                                                              int width, double* __restrict__ out) {
  __shared__ double red[256];
  const int c = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nbrBlocks; b += blockDim.x) s += partial[(int64_t)b * width + c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = red[0];
}

// C(j, m) for j <= nbrOrbitals, m <= holes, exact in 64 bits (Pascal's rule), on the device once per
// Hamiltonian; returns the number of (N-2)-subsets of the orbitals, or -1 when it exceeds 2^40
static int64_t adjoint_binomials(dh_hamiltonian* h, int holes, int* rc) {
  dh_context* c = h->ctx;
  const int n = h->basis->nbrOrbitals;
  std::vector<int64_t> B((size_t)(n + 1) * (holes + 1), 0);
  for (int j = 0; j <= n; ++j) {
    B[(size_t)j * (holes + 1)] = 1;
    for (int m = 1; m <= holes && j > 0; ++m) {
      const int64_t v = B[(size_t)(j - 1) * (holes + 1) + m - 1] + B[(size_t)(j - 1) * (holes + 1) + m];
      B[(size_t)j * (holes + 1) + m] = std::min<int64_t>(v, (int64_t)1 << 62);
    }
  }
  const int64_t total = B[(size_t)n * (holes + 1) + holes];
  *rc = DH_OK;
  if (total > ((int64_t)1 << 40)) return -1;
  if (!h->d_adjBinom) {
    if ((*rc = dh_alloc(c, (void**)&h->d_adjBinom, sizeof(int64_t) * B.size()))) return -1;
    hipError_t e = hipMemcpyAsync(h->d_adjBinom, B.data(), sizeof(int64_t) * B.size(), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      *rc = dh_fail(c, DH_E_HIP, hipGetErrorString(e));
      return -1;
    }
  }
  return total;
}

template <int NC, bool SAME>
static int adjoint_pass(dh_hamiltonian* h, int holes, int64_t subsets, int c0, const double* psi, const double* phi,
                        double* g) {
  dh_context* c = h->ctx;
  dh_basis* b = h->basis;
  size_t lds = sizeof(double) * h->nbrPairs * NC + sizeof(int64_t) * (b->nbrOrbitals + 1) * (holes + 1);
  if (lds > 150 * 1024) return dh_fail(c, DH_E_RANGE, "channel table exceeds LDS");
  auto kern = sphere_adjoint_kernel<NC, SAME>;
  if (lds > 64 * 1024) DH_HIP(c, hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int perCU = dh_blocks_per_cu(c, kern, 256, lds);
  int nb = (int)std::max<int64_t>(1, std::min<int64_t>((subsets + 255) / 256, (int64_t)c->nbrCU * std::max(perCU, 1)));
  double* scratch;
  int rc = dh_scratch(c, sizeof(double) * ((size_t)nb * NC + NC), &scratch);
  if (rc) return rc;
  hipLaunchKernelGGL(kern, dim3(nb), dim3(256), lds, c->stream, b->nbrOrbitals, holes, subsets, b->orbitalSum,
                     (const int64_t*)h->d_adjBinom, h->nbrChannels, c0, h->d_cg, h->nbrPairs, b->d_rankLo,
                     b->d_rankHi, b->rankBits, b->dim, psi, phi, scratch);
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(NC), dim3(256), 0, c->stream, scratch, nb, NC, scratch + (size_t)nb * NC);
  DH_HIP(c, hipGetLastError());
  double out[NC];
  DH_HIP(c, hipMemcpyAsync(out, scratch + (size_t)nb * NC, sizeof(double) * NC, hipMemcpyDeviceToHost, c->stream));
  DH_HIP(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < NC && c0 + i < h->nbrChannels; ++i) g[c0 + i] = out[i];
  return DH_OK;
}

extern "C" int dh_sphere_adjoint(dh_hamiltonian* h, const double* psi, const double* phi, double* g, double* hpsi) {
  if (!h || !psi || !phi || !g) return DH_E_ARG;
  if (h->kind != DH_HAM_SPHERE) return dh_fail(h->ctx, DH_E_ARG, "not a sphere Hamiltonian");
  if (!h->d_cg || h->nbrChannels <= 0) return dh_fail(h->ctx, DH_E_STATE, "channels not set");
  // the vector-Jacobian product H psi is one product of the same Hamiltonian (with its shift)
  if (hpsi) {
    int rc = dh_hv(h, psi, hpsi, 1, 0, 0, -1);
    if (rc) return rc;
  }
  const int holes = h->basis->nbrFermions - 2;
  if (holes < 0) {  // fewer than two particles: no two-body term
    DH_HIP(h->ctx, hipStreamSynchronize(h->ctx->stream));
    for (int i = 0; i < h->nbrChannels; ++i) g[i] = 0.0;
    return DH_OK;
  }
  int rc;
  const int64_t subsets = adjoint_binomials(h, holes, &rc);
  if (rc) return rc;
  if (subsets < 0) return dh_fail(h->ctx, DH_E_RANGE, "more than 2^40 (N-2)-particle states");
  // channels in groups of 16 (8 when they fit): three (or two) sums per channel stay in registers
  const bool same = psi == phi;
  for (int c0 = 0; c0 < h->nbrChannels; c0 += 16) {
    const int left = h->nbrChannels - c0;
    if (left <= 8)
      rc = same ? adjoint_pass<8, true>(h, holes, subsets, c0, psi, phi, g)
                : adjoint_pass<8, false>(h, holes, subsets, c0, psi, phi, g);
    else
      rc = same ? adjoint_pass<16, true>(h, holes, subsets, c0, psi, phi, g)
                : adjoint_pass<16, false>(h, holes, subsets, c0, psi, phi, g);
    if (rc) return rc;
  }
  return DH_OK;
}

// ----------------------------------------------------------------------------- stored elements
// AbstractQHEOnSphereHamiltonian::EnableFastMultiplication (FQHE/src/Hamiltonian/
// AbstractQHEOnSphereHamiltonian.cc:3814): the elements of every row are generated once into
// CSR arrays (64-bit row offsets, 32-bit columns, FP64 values) and each product then streams them.
// Gather form again: y[r] = sum_j val[j] x[col[j]] + shift x[r]; LPR lanes share a row and read
// its elements coalesced, then reduce by shuffles.
template <int K, int LPR>
__global__ void __launch_bounds__(512) csr_hv_kernel(const int64_t* __restrict__ rowPtr,
                                                     const int32_t* __restrict__ col,
                                                     const double* __restrict__ val, int64_t first, int64_t count,
                                                     const double* __restrict__ x, double* __restrict__ y, int add,
                                                     double shift) {
  const int lane = threadIdx.x % LPR;
  const int64_t rowsPerBlock = blockDim.x / LPR;
  const int64_t nbrTiles = (count + rowsPerBlock - 1) / rowsPerBlock;
  for (int64_t tile = dh_xcd_block(blockIdx.x, gridDim.x); tile < nbrTiles; tile += gridDim.x) {
    const int64_t local = tile * rowsPerBlock + threadIdx.x / LPR;
    const bool active = local < count;
    const int64_t row = first + (active ? local : 0);
    double acc[K];
#pragma unroll
    for (int v = 0; v < K; ++v) acc[v] = 0.0;
    if (active) {
      const int64_t end = rowPtr[row + 1];
      for (int64_t j = rowPtr[row] + lane; j < end; j += LPR) {
        const double c = val[j];
        const double* xr = x + (int64_t)col[j] * K;
        if constexpr (K % 2 == 0) {
#pragma unroll
          for (int v = 0; v < K; v += 2) {
            double2 p = *reinterpret_cast<const double2*>(xr + v);
            acc[v] += c * p.x;
            acc[v + 1] += c * p.y;
          }
        } else {
#pragma unroll
          for (int v = 0; v < K; ++v) acc[v] += c * xr[v];
        }
      }
    }
#pragma unroll
    for (int v = 0; v < K; ++v)
      for (int o = LPR / 2; o > 0; o >>= 1) acc[v] += __shfl_xor(acc[v], o, LPR);
    if (active && lane == 0) {
      const double* xr = x + row * K;
      double* yr = y + row * K;
#pragma unroll
      for (int v = 0; v < K; ++v) yr[v] = (add ? yr[v] : 0.0) + acc[v] + shift * xr[v];
    }
  }
}

template <int K>
static int csr_launch(dh_hamiltonian* h, const double* x, double* y, int add, int64_t first, int64_t count) {
  dh_context* c = h->ctx;
  constexpr int LPR = K >= 8 ? 16 : 32;
  auto kern = csr_hv_kernel<K, LPR>;
  const int perCU = dh_blocks_per_cu(c, kern, 512, 0);
  const int64_t tiles = (count + 512 / LPR - 1) / (512 / LPR);
  const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(tiles, (int64_t)c->nbrCU * perCU));
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(512), 0, c->stream, h->d_csrRow, h->d_csrCol, h->d_csrVal,
                     first, count, x, y, add, h->shift);
  DH_HIP(c, hipGetLastError());
  return DH_OK;
}

int dh_csr_hv(dh_hamiltonian* h, const double* x, double* y, int k, int add, int64_t first, int64_t count) {
  switch (k) {
    case 1: return csr_launch<1>(h, x, y, add, first, count);
    case 2: return csr_launch<2>(h, x, y, add, first, count);
    case 4: return csr_launch<4>(h, x, y, add, first, count);
    case 8: return csr_launch<8>(h, x, y, add, first, count);
    case 16: return csr_launch<16>(h, x, y, add, first, count);
    default: return dh_fail(h->ctx, DH_E_ARG, "k must be 1, 2, 4, 8 or 16");
  }
}

void dh_csr_free(dh_hamiltonian* h) {
  for (void* p : {(void*)h->d_csrRow, (void*)h->d_csrCol, (void*)h->d_csrVal})
    if (p) hipFree(p);
  h->d_csrRow = nullptr;
  h->d_csrCol = nullptr;
  h->d_csrVal = nullptr;
  h->csrNnz = 0;
}

extern "C" int dh_hamiltonian_release_precomputed(dh_hamiltonian* h) {
  if (!h) return DH_E_ARG;
  hipStreamSynchronize(h->ctx->stream);
  dh_csr_free(h);
  return DH_OK;
}

extern "C" int dh_hamiltonian_precompute(dh_hamiltonian* h, int64_t maxBytes, int64_t* bytes) {
  if (!h) return DH_E_ARG;
  dh_context* c = h->ctx;
  if (h->kind != DH_HAM_SPHERE) return dh_fail(c, DH_E_ARG, "stored elements are implemented for two-body sphere Hamiltonians");
  const dh_basis* b = h->basis;
  const int64_t dim = b->dim;
  if (dim >= (int64_t(1) << 31)) return dh_fail(c, DH_E_RANGE, "stored elements need fewer than 2^31 rows");
  hipStreamSynchronize(c->stream);
  dh_csr_free(h);
  // element count per row, offsets by a host prefix sum
  std::vector<int64_t> rowPtr(dim + 1, 0);
  const int64_t chunk = 1 << 24;
  int64_t* d_counts = nullptr;
  int rc;
  if ((rc = dh_alloc(c, (void**)&d_counts, sizeof(int64_t) * std::max<int64_t>(1, std::min(chunk, dim))))) return rc;
  for (int64_t r0 = 0; r0 < dim; r0 += chunk) {
    const int64_t r1 = std::min(dim, r0 + chunk);
    hipLaunchKernelGGL(fs_elements_kernel, dim3((unsigned)((r1 - r0 + 255) / 256)), dim3(256), 0, c->stream,
                       b->d_states, dim, b->nbrOrbitals, h->d_factors, h->d_pairOffset, b->d_rankLo, b->d_rankHi,
                       b->rankBits, r0, r1, d_counts, (int64_t*)nullptr, (double*)nullptr, (int64_t)0,
                       (const int64_t*)nullptr, (int32_t*)nullptr);
    hipMemcpyAsync(rowPtr.data() + 1 + r0, d_counts, sizeof(int64_t) * (r1 - r0), hipMemcpyDeviceToHost, c->stream);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      hipFree(d_counts);
      return dh_fail(c, DH_E_HIP, hipGetErrorString(e));
    }
  }
  hipFree(d_counts);
  for (int64_t r = 0; r < dim; ++r) rowPtr[r + 1] += rowPtr[r];
  const int64_t nnz = rowPtr[dim];
  const int64_t need = (int64_t)sizeof(int64_t) * (dim + 1) + nnz * (int64_t)(sizeof(int32_t) + sizeof(double));
  if (bytes) *bytes = need;
  if (need > maxBytes) return dh_fail(c, DH_E_NOMEM, "the stored elements need more than max_bytes");
  if ((rc = dh_alloc(c, (void**)&h->d_csrRow, sizeof(int64_t) * (dim + 1))) ||
      (rc = dh_alloc(c, (void**)&h->d_csrCol, sizeof(int32_t) * std::max<int64_t>(1, nnz))) ||
      (rc = dh_alloc(c, (void**)&h->d_csrVal, sizeof(double) * std::max<int64_t>(1, nnz)))) {
    dh_csr_free(h);
    return rc;
  }
  DH_HIP(c, hipMemcpyAsync(h->d_csrRow, rowPtr.data(), sizeof(int64_t) * (dim + 1), hipMemcpyHostToDevice, c->stream));
  for (int64_t r0 = 0; r0 < dim; r0 += chunk) {
    const int64_t r1 = std::min(dim, r0 + chunk);
    hipLaunchKernelGGL(fs_elements_kernel, dim3((unsigned)((r1 - r0 + 255) / 256)), dim3(256), 0, c->stream,
                       b->d_states, dim, b->nbrOrbitals, h->d_factors, h->d_pairOffset, b->d_rankLo, b->d_rankHi,
                       b->rankBits, r0, r1, (int64_t*)nullptr, (int64_t*)nullptr, h->d_csrVal, (int64_t)0,
                       (const int64_t*)h->d_csrRow, h->d_csrCol);
  }
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    dh_csr_free(h);
    return dh_fail(c, DH_E_HIP, hipGetErrorString(e));
  }
  h->csrNnz = nnz;
  return DH_OK;
}
