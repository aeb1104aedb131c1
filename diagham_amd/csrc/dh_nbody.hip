// Three-body interactions on the sphere (gather form) for gfx950.
//
// Reference loop replaced: AbstractQHEOnSphereNBodyInteractionHamiltonian::LowLevelAddMultiply
// (FQHE/src/Hamiltonian/AbstractQHEOnSphereNBodyInteractionHamiltonian.cc:107) through
// EvaluateMNNBodyAddMultiplyComponent with explicit MN factors (.h:259-283), i.e.
//   H = sum_S sum_{i,l} F_S[i][l] ProdAd(M_l) ProdA(N_i)
// with FermionOnSphere::ProdA (FermionOnSphere.cc:716) and ProdAd (:875) inlined.  The tuples
// N_i, M_l of orbital sum S are given by the caller (descending orbitals, as the reference's
// GetAllSkewSymmetricIndices lists them); F_S is symmetric, so row r of H is the list of terms
// generated from state r and y[r] is gathered, one row per LPR lanes, without atomics.
//
// Per row: every triple of occupied orbitals n0 > n1 > n2 is annihilated (sign: occupied
// orbitals at and above each, rightmost first), its position i in the list of sum S comes from a
// table indexed by the combinatorial rank of the triple, and every listed tuple of the same sum
// whose orbitals are free is created (sign: occupied orbitals above each, rightmost first).
// Tuple masks, positions and per-sum offsets are staged in LDS; the factor matrices stay in
// global memory (L2).
#include <algorithm>
#include <cstring>
#include <vector>

#include "dh_index.h"
#include "dh_internal.h"
#include "dh_sphere_row.h"

struct NBodyArgs {
  const uint64_t* states;
  int64_t dim;
  int nbrTuples, nbrSums;
  int64_t posSize;
  const uint64_t* tupleMask;
  const int* tuplePos;
  const int* sumTupleOff;
  const int64_t* sumFactorOff;
  const double* factors;
  const uint32_t* rankLo;
  const uint32_t* rankHi;
  int rankBits;
  double shift;
};

__device__ __forceinline__ int64_t dh_triple_rank(int n0, int n1, int n2) {
  // C(n0, 3) + C(n1, 2) + C(n2, 1) for n0 > n1 > n2
  return (int64_t)n0 * (n0 - 1) * (n0 - 2) / 6 + (int64_t)n1 * (n1 - 1) / 2 + n2;
}

__device__ __forceinline__ void nbody_stage(const NBodyArgs& a, uint64_t*& sMask, int*& sPos, int*& sOff,
                                            int64_t*& sFOff) {
  extern __shared__ uint64_t smemNB[];
  sMask = smemNB;
  sFOff = reinterpret_cast<int64_t*>(sMask + a.nbrTuples);
  sPos = reinterpret_cast<int*>(sFOff + a.nbrSums + 1);
  sOff = sPos + a.posSize;
  for (int i = threadIdx.x; i < a.nbrTuples; i += blockDim.x) sMask[i] = a.tupleMask[i];
  for (int i = threadIdx.x; i <= a.nbrSums; i += blockDim.x) {
    sFOff[i] = a.sumFactorOff[i];
    sOff[i] = a.sumTupleOff[i];
  }
  for (int64_t i = threadIdx.x; i < a.posSize; i += blockDim.x) sPos[i] = a.tuplePos[i];
  __syncthreads();
}

template <int K, int LPR>
__global__ void __launch_bounds__(1024) nbody3_hv_kernel(NBodyArgs a, const double* __restrict__ x,
                                                         double* __restrict__ y, int add, int64_t first,
                                                         int64_t count) {
  uint64_t* sMask;
  int *sPos, *sOff;
  int64_t* sFOff;
  nbody_stage(a, sMask, sPos, sOff, sFOff);
  const int lane = threadIdx.x % LPR;
  const int64_t rowsPerBlock = blockDim.x / LPR;
  const int64_t nbrTiles = (count + rowsPerBlock - 1) / rowsPerBlock;
  for (int64_t tile = dh_xcd_block(blockIdx.x, gridDim.x); tile < nbrTiles; tile += gridDim.x) {
    const int64_t local = tile * rowsPerBlock + threadIdx.x / LPR;
    const bool active = local < count;
    const int64_t row = first + (active ? local : 0);
    const uint64_t s = active ? a.states[row] : 0ull;
    double acc[K];
#pragma unroll
    for (int v = 0; v < K; ++v) acc[v] = 0.0;
    for (uint64_t r0 = s; r0; r0 &= r0 - 1) {
      const int n0 = __builtin_ctzll(r0);
      for (uint64_t r1 = s & ((1ull << n0) - 1ull); r1; r1 &= r1 - 1) {
        const int n1 = __builtin_ctzll(r1);
        for (uint64_t r2 = s & ((1ull << n1) - 1ull); r2; r2 &= r2 - 1) {
          const int n2 = __builtin_ctzll(r2);
          const int i = sPos[dh_triple_rank(n0, n1, n2)];
          if (i < 0) continue;
          const uint64_t s1 = s ^ (1ull << n2), s2 = s1 ^ (1ull << n1), t = s2 ^ (1ull << n0);
          const int pA = __popcll(s >> n2) + __popcll(s1 >> n1) + __popcll(s2 >> n0);
          const int S = n0 + n1 + n2;
          const int base = sOff[S], cnt = sOff[S + 1] - base;
          const double* F = a.factors + sFOff[S] + (int64_t)i * cnt;
          for (int l = lane; l < cnt; l += LPR) {
            const uint64_t m = sMask[base + l];
            if (t & m) continue;
            const uint64_t mUp = m & (m - 1);
            const int m2 = __builtin_ctzll(m), m1 = __builtin_ctzll(mUp), m0 = __builtin_ctzll(mUp & (mUp - 1));
            const uint64_t t1 = t | (1ull << m2), t2 = t1 | (1ull << m1);
            const int pC = __popcll(t >> m2) + __popcll(t1 >> m1) + __popcll(t2 >> m0);
            double coef = F[l];
            if ((pA + pC) & 1) coef = -coef;
            const int64_t idx = dh_find_rank(t | m, a.rankLo, a.rankHi, a.rankBits, a.dim);
            dh_gather_fma<K>(acc, x, idx, coef);
          }
        }
      }
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

// matrix elements per row (every created tuple of the same sum with free orbitals)
__global__ void nbody3_count_kernel(NBodyArgs a, int64_t rowBegin, int64_t rowEnd, int64_t* __restrict__ counts) {
  uint64_t* sMask;
  int *sPos, *sOff;
  int64_t* sFOff;
  nbody_stage(a, sMask, sPos, sOff, sFOff);
  const int64_t row = rowBegin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= rowEnd) return;
  const uint64_t s = a.states[row];
  int64_t n = 0;
  for (uint64_t r0 = s; r0; r0 &= r0 - 1) {
    const int n0 = __builtin_ctzll(r0);
    for (uint64_t r1 = s & ((1ull << n0) - 1ull); r1; r1 &= r1 - 1) {
      const int n1 = __builtin_ctzll(r1);
      for (uint64_t r2 = s & ((1ull << n1) - 1ull); r2; r2 &= r2 - 1) {
        const int n2 = __builtin_ctzll(r2);
        if (sPos[dh_triple_rank(n0, n1, n2)] < 0) continue;
        const uint64_t t = s ^ (1ull << n0) ^ (1ull << n1) ^ (1ull << n2);
        const int S = n0 + n1 + n2;
        for (int l = sOff[S]; l < sOff[S + 1]; ++l) n += (t & sMask[l]) ? 0 : 1;
      }
    }
  }
  counts[row - rowBegin] = n;
}

static NBodyArgs nbody_args(const dh_hamiltonian* h) {
  const dh_basis* b = h->basis;
  NBodyArgs a;
  a.states = b->d_states;
  a.dim = b->dim;
  a.nbrTuples = h->nbrTuples;
  a.nbrSums = h->nbrTupleSums;
  a.posSize = h->tuplePosSize;
  a.tupleMask = h->d_tupleMask;
  a.tuplePos = h->d_tuplePos;
  a.sumTupleOff = h->d_sumTupleOff;
  a.sumFactorOff = h->d_sumFactorOff;
  a.factors = h->d_nbodyFactors;
  a.rankLo = b->d_rankLo;
  a.rankHi = b->d_rankHi;
  a.rankBits = b->rankBits;
  a.shift = h->shift;
  return a;
}

static size_t nbody_lds(const dh_hamiltonian* h) {
  return sizeof(uint64_t) * h->nbrTuples + sizeof(int64_t) * (h->nbrTupleSums + 1) +
         sizeof(int) * (h->tuplePosSize + h->nbrTupleSums + 1);
}

template <int K, int LPR>
static int nbody_launch(dh_hamiltonian* h, const double* x, double* y, int add, int64_t first, int64_t count) {
  dh_context* c = h->ctx;
  const size_t lds = nbody_lds(h);
  auto kern = nbody3_hv_kernel<K, LPR>;
  if (lds > 64 * 1024)
    DH_HIP(c, hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int perCU = dh_blocks_per_cu(c, kern, h->blockSize, lds);
  const int64_t rowsPerBlock = h->blockSize / LPR;
  const int64_t tiles = (count + rowsPerBlock - 1) / rowsPerBlock;
  const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(tiles, (int64_t)c->nbrCU * perCU));
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(h->blockSize), lds, c->stream, nbody_args(h), x, y, add, first,
                     count);
  DH_HIP(c, hipGetLastError());
  return DH_OK;
}

template <int K>
static int nbody_lpr(dh_hamiltonian* h, const double* x, double* y, int add, int64_t first, int64_t count) {
  switch (h->lanesPerRow) {
    case 1: return nbody_launch<K, 1>(h, x, y, add, first, count);
    case 2: return nbody_launch<K, 2>(h, x, y, add, first, count);
    case 4: return nbody_launch<K, 4>(h, x, y, add, first, count);
    default: return nbody_launch<K, 8>(h, x, y, add, first, count);
  }
}

int dh_nbody_hv(dh_hamiltonian* h, const double* x, double* y, int k, int add, int64_t first, int64_t count) {
  switch (k) {
    case 1: return nbody_lpr<1>(h, x, y, add, first, count);
    case 2: return nbody_lpr<2>(h, x, y, add, first, count);
    case 4: return nbody_lpr<4>(h, x, y, add, first, count);
    case 8: return nbody_lpr<8>(h, x, y, add, first, count);
    case 16: return nbody_lpr<16>(h, x, y, add, first, count);
    default: return dh_fail(h->ctx, DH_E_ARG, "k must be 1, 2, 4, 8 or 16");
  }
}

int dh_nbody_count_elements(dh_hamiltonian* h, int64_t* nnz) {
  dh_context* c = h->ctx;
  const dh_basis* b = h->basis;
  const int64_t chunk = 1 << 22;
  int64_t* d_counts = nullptr;
  int rc = dh_alloc(c, (void**)&d_counts, sizeof(int64_t) * std::max<int64_t>(1, std::min(chunk, b->dim)));
  if (rc) return rc;
  const size_t lds = nbody_lds(h);
  if (lds > 64 * 1024)
    DH_HIP(c, hipFuncSetAttribute((const void*)nbody3_count_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds));
  std::vector<int64_t> hc;
  int64_t total = 0;
  for (int64_t r0 = 0; r0 < b->dim; r0 += chunk) {
    const int64_t r1 = std::min(b->dim, r0 + chunk);
    hipLaunchKernelGGL(nbody3_count_kernel, dim3((unsigned)((r1 - r0 + 255) / 256)), dim3(256), lds, c->stream,
                       nbody_args(h), r0, r1, d_counts);
    hc.resize(r1 - r0);
    hipMemcpyAsync(hc.data(), d_counts, sizeof(int64_t) * (r1 - r0), hipMemcpyDeviceToHost, c->stream);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      hipFree(d_counts);
      return dh_fail(c, DH_E_HIP, hipGetErrorString(e));
    }
    for (int64_t v : hc) total += v;
  }
  hipFree(d_counts);
  *nnz = total;
  return DH_OK;
}

void dh_nbody_free(dh_hamiltonian* h) {
  for (void* p : {(void*)h->d_tupleMask, (void*)h->d_tuplePos, (void*)h->d_sumTupleOff, (void*)h->d_sumFactorOff,
                  (void*)h->d_nbodyFactors})
    if (p) hipFree(p);
}

extern "C" int dh_sphere_nbody_hamiltonian_create(dh_context* c, dh_hamiltonian** out, dh_basis* b, int nbody,
                                                  const int* sumCount, int nbrSums, const int* tuples,
                                                  const double* factors) {
  if (!c || !out || !b || !sumCount || !tuples || !factors || nbrSums <= 0) return DH_E_ARG;
  *out = nullptr;
  if (c->closed) return dh_fail(c, DH_E_STATE, "context destroyed");
  if (b->kind != DH_BASIS_FERMION_SPHERE) return dh_fail(c, DH_E_ARG, "basis is not a fermion-on-sphere basis");
  if (nbody != 3) return dh_fail(c, DH_E_ARG, "the n-body kernel handles three-body terms (nbody = 3)");
  const int nOrb = b->nbrOrbitals;
  if (nbrSums > 3 * (nOrb - 1) + 1) return dh_fail(c, DH_E_ARG, "more orbital sums than the basis allows");
  const int64_t posSize = (int64_t)nOrb * (nOrb - 1) * (nOrb - 2) / 6;
  std::vector<int> pos(std::max<int64_t>(posSize, 1), -1), tupleOff(nbrSums + 1, 0);
  std::vector<int64_t> factorOff(nbrSums + 1, 0);
  std::vector<uint64_t> masks;
  int64_t t = 0;
  for (int S = 0; S < nbrSums; ++S) {
    const int cnt = sumCount[S];
    if (cnt < 0) return dh_fail(c, DH_E_ARG, "negative tuple count");
    tupleOff[S + 1] = tupleOff[S] + cnt;
    factorOff[S + 1] = factorOff[S] + (int64_t)cnt * cnt;
    for (int i = 0; i < cnt; ++i, ++t) {
      const int n0 = tuples[3 * t], n1 = tuples[3 * t + 1], n2 = tuples[3 * t + 2];
      if (!(n0 > n1 && n1 > n2 && n2 >= 0 && n0 < nOrb))
        return dh_fail(c, DH_E_ARG, "tuples must hold three distinct orbitals in descending order");
      if (n0 + n1 + n2 != S) return dh_fail(c, DH_E_ARG, "a tuple is listed under the wrong orbital sum");
      const int64_t r = (int64_t)n0 * (n0 - 1) * (n0 - 2) / 6 + (int64_t)n1 * (n1 - 1) / 2 + n2;
      if (pos[r] >= 0) return dh_fail(c, DH_E_ARG, "a tuple is listed twice");
      pos[r] = i;
      masks.push_back((1ull << n0) | (1ull << n1) | (1ull << n2));
    }
  }
  // the gather form reads row i of F_S for the annihilated tuple i: F_S must be symmetric
  double scale = 0.0, asym = 0.0;
  for (int S = 0; S < nbrSums; ++S) {
    const int cnt = sumCount[S];
    const double* F = factors + factorOff[S];
    for (int i = 0; i < cnt; ++i)
      for (int l = 0; l < cnt; ++l) {
        scale = std::max(scale, std::fabs(F[(int64_t)i * cnt + l]));
        asym = std::max(asym, std::fabs(F[(int64_t)i * cnt + l] - F[(int64_t)l * cnt + i]));
      }
  }
  if (asym > 1e-12 * scale) return dh_fail(c, DH_E_ARG, "the factor matrices are not symmetric");
  dh_hamiltonian* h = new dh_hamiltonian();
  h->ctx = c;
  h->basis = b;
  h->kind = DH_HAM_SPHERE_NBODY;
  h->nbody = nbody;
  h->nbrTuples = (int)masks.size();
  h->nbrTupleSums = nbrSums;
  h->tuplePosSize = posSize;
  h->nbrNBodyFactors = factorOff[nbrSums];
  if (nbody_lds(h) > 160 * 1024) {
    delete h;
    return dh_fail(c, DH_E_RANGE, "the tuple tables exceed LDS");
  }
  int rc;
  if ((rc = dh_alloc(c, (void**)&h->d_tupleMask, sizeof(uint64_t) * std::max<size_t>(1, masks.size()))) ||
      (rc = dh_alloc(c, (void**)&h->d_tuplePos, sizeof(int) * pos.size())) ||
      (rc = dh_alloc(c, (void**)&h->d_sumTupleOff, sizeof(int) * tupleOff.size())) ||
      (rc = dh_alloc(c, (void**)&h->d_sumFactorOff, sizeof(int64_t) * factorOff.size())) ||
      (rc = dh_alloc(c, (void**)&h->d_nbodyFactors, sizeof(double) * std::max<int64_t>(1, factorOff[nbrSums])))) {
    dh_hamiltonian_destroy(h);
    return rc;
  }
  DH_HIP(c, hipMemcpyAsync(h->d_tupleMask, masks.data(), sizeof(uint64_t) * masks.size(), hipMemcpyHostToDevice,
                           c->stream));
  DH_HIP(c, hipMemcpyAsync(h->d_tuplePos, pos.data(), sizeof(int) * pos.size(), hipMemcpyHostToDevice, c->stream));
  DH_HIP(c, hipMemcpyAsync(h->d_sumTupleOff, tupleOff.data(), sizeof(int) * tupleOff.size(), hipMemcpyHostToDevice,
                           c->stream));
  DH_HIP(c, hipMemcpyAsync(h->d_sumFactorOff, factorOff.data(), sizeof(int64_t) * factorOff.size(),
                           hipMemcpyHostToDevice, c->stream));
  DH_HIP(c, hipMemcpyAsync(h->d_nbodyFactors, factors, sizeof(double) * factorOff[nbrSums], hipMemcpyHostToDevice,
                           c->stream));
  DH_HIP(c, hipStreamSynchronize(c->stream));
  h->lanesPerRow = (b->dim < (int64_t)c->nbrCU * 256) ? 4 : 1;
  h->blockSize = 512;
  h->indexMode = DH_INDEX_RANK;
  *out = h;
  return DH_OK;
}
