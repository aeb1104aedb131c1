// Bases of the on-the-fly product: fermions on the sphere and spin-1/2 chains.
//
// FermionOnSphere (FQHE/src/HilbertSpace/FermionOnSphere.cc:76) orders its states as
// descending unsigned 64-bit integers (GenerateStates :1621 places the highest orbital
// first).  Here the basis is produced on the device by unranking: with
//   T[j][n][s] = number of n-subsets of orbitals {0..j-1} whose indices sum to s,
// the state of index i is built from the top orbital down, taking orbital j when i falls
// among the T[j][n-1][s-j] states that occupy it.  The same table gives the
// combinatorial rank used as FindStateIndex (:1431) in DH_INDEX_RANK mode:
//   index(state) = dim - 1 - sum_{occupied j} T[j][n_j][s_j]
// (n_j, s_j = number and index sum of the occupied orbitals <= j), split into a low-bit
// table and a high-bit table so one lookup costs two table reads.
//
// Spin1_2Chain (Spin/src/HilbertSpace/Spin1_2Chain.cc:113, :331) decides site 0 first and
// keeps bits >= L at 1; its order is descending in the bit-reversed L-bit word, and
//   index(state) = dim - 1 - sum_{up sites i} C(L-1-i, K_i),  K_i = up spins at sites >= i.
#include <algorithm>
#include <cstring>

#include "dh_internal.h"

// ----------------------------------------------------------------------------- host tables
struct SubsetCountTable {
  int nbrOrbitals = 0, maxN = 0, maxSum = 0;
  std::vector<int64_t> t;  // [(j*(maxN+1) + n)*(maxSum+1) + s]
  int64_t at(int j, int n, int s) const {
    if (j < 0 || j > nbrOrbitals || n < 0 || n > maxN || s < 0 || s > maxSum) return 0;
    return t[((size_t)j * (maxN + 1) + n) * (maxSum + 1) + s];
  }
};

static SubsetCountTable make_subset_table(int nbrOrbitals, int maxN) {
  SubsetCountTable T;
  T.nbrOrbitals = nbrOrbitals;
  T.maxN = maxN;
  int maxSum = 0;
  for (int i = 0; i < maxN; ++i) maxSum += nbrOrbitals - 1 - i;
  T.maxSum = std::max(maxSum, 0);
  T.t.assign((size_t)(nbrOrbitals + 1) * (maxN + 1) * (T.maxSum + 1), 0);
  auto idx = [&](int j, int n, int s) { return ((size_t)j * (maxN + 1) + n) * (T.maxSum + 1) + s; };
  T.t[idx(0, 0, 0)] = 1;
  for (int j = 0; j < nbrOrbitals; ++j)
    for (int n = 0; n <= maxN; ++n)
      for (int s = 0; s <= T.maxSum; ++s) {
        int64_t v = T.t[idx(j, n, s)];
        if (n >= 1 && s >= j) v += T.t[idx(j, n - 1, s - j)];
        T.t[idx(j + 1, n, s)] = v;
      }
  return T;
}

extern "C" int64_t dh_fermion_sphere_dimension(int nbrFermions, int totalLz, int lzMax) {
  if (nbrFermions <= 0 || lzMax < 0 || lzMax > 62 || nbrFermions > lzMax + 1) return 0;
  int twiceSum = totalLz + nbrFermions * lzMax;
  if (twiceSum < 0 || (twiceSum & 1)) return 0;
  SubsetCountTable T = make_subset_table(lzMax + 1, nbrFermions);
  return T.at(lzMax + 1, nbrFermions, twiceSum >> 1);
}

// ----------------------------------------------------------------------------- kernels
// unrank fermion states: descending order, table T in global memory
__global__ void fs_unrank_kernel(uint64_t* __restrict__ states, int64_t dim, int nbrOrbitals, int nbrFermions,
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

// lut[v] = number of states whose key (= state >> shift) is >= v, for v in [0, 2^bits];
// keys are non-increasing along the descending basis
__global__ void fs_lut_kernel(const uint64_t* __restrict__ states, int64_t dim, int shift, int bits,
                              int* __restrict__ lut) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t size = (1ll << bits);
  if (v > size) return;
  // first index whose key < v
  int64_t lo = 0, hi = dim;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if ((int64_t)(states[mid] >> shift) >= v)
      lo = mid + 1;
    else
      hi = mid;
  }
  lut[v] = (int)lo;
}

// spin chain: unrank in bit-reversed descending order with binomials
__global__ void spin_unrank_kernel(uint32_t* __restrict__ words, int64_t dim, int L, int nbrUp,
                                   const uint32_t* __restrict__ binom) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dim) return;
  int64_t rem = i;
  int n = nbrUp;
  uint32_t downMask = 0;
  for (int site = 0; site < L; ++site) {
    // states with this site up come first: C(L-1-site, n-1) of them
    int64_t c = (n > 0) ? (int64_t)binom[(L - 1 - site) * 33 + (n - 1)] : 0;
    if (rem < c) {
      --n;
    } else {
      rem -= c;
      downMask |= (1u << site);
    }
  }
  words[i] = 0xffffffffu & ~downMask;
}

__device__ __forceinline__ uint32_t spin_key(uint32_t word, int L) {
  uint32_t s = (L == 32) ? word : (word & ((1u << L) - 1u));
  return __builtin_bitreverse32(s) >> (32 - L);
}

__global__ void spin_lut_kernel(const uint32_t* __restrict__ words, int64_t dim, int L, int shift, int bits,
                                int* __restrict__ lut) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t size = (1ll << bits);
  if (v > size) return;
  int64_t lo = 0, hi = dim;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if ((int64_t)(spin_key(words[mid], L) >> shift) >= v)
      lo = mid + 1;
    else
      hi = mid;
  }
  lut[v] = (int)lo;
}

// ----------------------------------------------------------------------------- basis creation
static int upload(dh_context* c, void** d, const void* h, size_t bytes) {
  int rc = dh_alloc(c, d, bytes);
  if (rc) return rc;
  DH_HIP(c, hipMemcpyAsync(*d, h, bytes, hipMemcpyHostToDevice, c->stream));
  DH_HIP(c, hipStreamSynchronize(c->stream));
  return DH_OK;
}

// rank tables for a fermion sector (host): lo over the low `bits` orbitals, hi over the rest
static void fermion_rank_tables(const SubsetCountTable& T, int nbrOrbitals, int N, int S, int bits,
                                std::vector<uint32_t>& lo, std::vector<uint32_t>& hi) {
  lo.assign((size_t)1 << bits, 0);
  for (uint64_t l = 0; l < lo.size(); ++l) {
    int n = 0, s = 0;
    int64_t acc = 0;
    for (uint64_t r = l; r; r &= r - 1) {
      int j = __builtin_ctzll(r);
      ++n;
      s += j;
      acc += T.at(j, n, s);
    }
    lo[l] = (uint32_t)acc;
  }
  int hb = nbrOrbitals - bits;
  hi.assign((size_t)1 << hb, 0);
  for (uint64_t h = 0; h < hi.size(); ++h) {
    int popH = __builtin_popcountll(h);
    int sumH = 0;
    for (uint64_t r = h; r; r &= r - 1) sumH += bits + __builtin_ctzll(r);
    int n = N - popH, s = S - sumH;
    if (n < 0 || s < 0) continue;
    int64_t acc = 0;
    for (uint64_t r = h; r; r &= r - 1) {
      int j = bits + __builtin_ctzll(r);
      ++n;
      s += j;
      acc += T.at(j, n, s);
    }
    hi[h] = (uint32_t)acc;
  }
}

extern "C" int dh_fermion_sphere_create(dh_context* c, dh_basis** out, int N, int totalLz, int lzMax) {
  if (!c || !out) return DH_E_ARG;
  *out = nullptr;
  if (c->closed) return dh_fail(c, DH_E_STATE, "context destroyed");
  if (N <= 0 || lzMax < 0 || N > lzMax + 1) return dh_fail(c, DH_E_ARG, "need 0 < nbr_fermions <= lz_max + 1");
  if (lzMax > 62) return dh_fail(c, DH_E_RANGE, "lz_max must be <= 62 (64-bit states)");
  int twiceSum = totalLz + N * lzMax;
  if (twiceSum < 0 || (twiceSum & 1)) return dh_fail(c, DH_E_ARG, "total_lz has the wrong parity or range");
  hipSetDevice(c->device);
  int nbrOrbitals = lzMax + 1;
  SubsetCountTable T = make_subset_table(nbrOrbitals, N);
  int S = twiceSum >> 1;
  int64_t dim = T.at(nbrOrbitals, N, S);
  if (dim <= 0) return dh_fail(c, DH_E_ARG, "empty Hilbert space");
  if (dim >= (1ll << 31)) return dh_fail(c, DH_E_RANGE, "sector dimension must be < 2^31");
  dh_basis* b = new dh_basis();
  b->ctx = c;
  b->kind = DH_BASIS_FERMION_SPHERE;
  b->dim = dim;
  b->nbrFermions = N;
  b->lzMax = lzMax;
  b->totalLz = totalLz;
  b->nbrOrbitals = nbrOrbitals;
  b->orbitalSum = S;
  int rc;
  int64_t* d_T = nullptr;
  if ((rc = upload(c, (void**)&d_T, T.t.data(), sizeof(int64_t) * T.t.size()))) {
    delete b;
    return rc;
  }
  if ((rc = dh_alloc(c, (void**)&b->d_states, sizeof(uint64_t) * dim))) {
    hipFree(d_T);
    delete b;
    return rc;
  }
  hipLaunchKernelGGL(fs_unrank_kernel, dim3((unsigned)((dim + 255) / 256)), dim3(256), 0, c->stream, b->d_states,
                     dim, nbrOrbitals, N, S, d_T, T.maxN, T.maxSum);
  // LUT over the top bits (binary-search mode)
  b->lutBits = std::min(nbrOrbitals, 14);
  b->lutShift = nbrOrbitals - b->lutBits;
  if ((rc = dh_alloc(c, (void**)&b->d_lut, sizeof(int) * (((size_t)1 << b->lutBits) + 1)))) {
    hipFree(d_T);
    dh_basis_destroy(b);
    return rc;
  }
  hipLaunchKernelGGL(fs_lut_kernel, dim3((unsigned)(((1 << b->lutBits) + 1 + 255) / 256)), dim3(256), 0, c->stream,
                     b->d_states, dim, b->lutShift, b->lutBits, b->d_lut);
  // rank tables
  b->rankBits = dh_rank_split(nbrOrbitals);
  std::vector<uint32_t> lo, hi;
  fermion_rank_tables(T, nbrOrbitals, N, S, b->rankBits, lo, hi);
  b->rankLoSize = (int64_t)lo.size();
  b->rankHiSize = (int64_t)hi.size();
  const std::vector<uint16_t> lo16 = dh_rank_lo16(lo);
  if ((rc = upload(c, (void**)&b->d_rankLo, lo.data(), sizeof(uint32_t) * lo.size())) ||
      (rc = upload(c, (void**)&b->d_rankHi, hi.data(), sizeof(uint32_t) * hi.size())) ||
      (!lo16.empty() && (rc = upload(c, (void**)&b->d_rankLo16, lo16.data(), sizeof(uint16_t) * lo16.size())))) {
    hipFree(d_T);
    dh_basis_destroy(b);
    return rc;
  }
  hipError_t e = hipStreamSynchronize(c->stream);
  hipFree(d_T);
  if (e != hipSuccess) {
    dh_basis_destroy(b);
    return dh_fail(c, DH_E_HIP, std::string("basis generation: ") + hipGetErrorString(e));
  }
  *out = b;
  return DH_OK;
}

extern "C" int dh_spin_chain_create(dh_context* c, dh_basis** out, int L, int sz) {
  if (!c || !out) return DH_E_ARG;
  *out = nullptr;
  if (c->closed) return dh_fail(c, DH_E_STATE, "context destroyed");
  if (L < 2 || L > 32) return dh_fail(c, DH_E_RANGE, "chain_length must be in [2, 32]");
  if (((L + sz) & 1) || sz > L || sz < -L) return dh_fail(c, DH_E_ARG, "sz inconsistent with chain_length");
  hipSetDevice(c->device);
  int nbrUp = (L + sz) / 2;
  int64_t dim = (int64_t)dh_binom_host(L, nbrUp);
  if (dim >= (1ll << 31)) return dh_fail(c, DH_E_RANGE, "sector dimension must be < 2^31");
  dh_basis* b = new dh_basis();
  b->ctx = c;
  b->kind = DH_BASIS_SPIN_CHAIN;
  b->dim = dim;
  b->chainLength = L;
  b->nbrUp = nbrUp;
  b->sz = sz;
  std::vector<uint32_t> binom(33 * 33, 0);
  for (int n = 0; n <= 32; ++n)
    for (int k = 0; k <= 32; ++k) binom[n * 33 + k] = (uint32_t)std::min<uint64_t>(dh_binom_host(n, k), 0xffffffffu);
  int rc;
  if ((rc = upload(c, (void**)&b->d_binom, binom.data(), sizeof(uint32_t) * binom.size())) ||
      (rc = dh_alloc(c, (void**)&b->d_words, sizeof(uint32_t) * dim))) {
    dh_basis_destroy(b);
    return rc;
  }
  hipLaunchKernelGGL(spin_unrank_kernel, dim3((unsigned)((dim + 255) / 256)), dim3(256), 0, c->stream, b->d_words,
                     dim, L, nbrUp, b->d_binom);
  b->lutBits = std::min(L, 14);
  b->lutShift = L - b->lutBits;
  if ((rc = dh_alloc(c, (void**)&b->d_lut, sizeof(int) * (((size_t)1 << b->lutBits) + 1)))) {
    dh_basis_destroy(b);
    return rc;
  }
  hipLaunchKernelGGL(spin_lut_kernel, dim3((unsigned)(((1 << b->lutBits) + 1 + 255) / 256)), dim3(256), 0,
                     c->stream, b->d_words, dim, L, b->lutShift, b->lutBits, b->d_lut);
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    dh_basis_destroy(b);
    return dh_fail(c, DH_E_HIP, std::string("basis generation: ") + hipGetErrorString(e));
  }
  *out = b;
  return DH_OK;
}

extern "C" int dh_basis_destroy(dh_basis* b) {
  if (!b) return DH_E_ARG;
  dh_context* c = b->ctx;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  if (b->d_states) hipFree(b->d_states);
  if (b->d_words) hipFree(b->d_words);
  if (b->d_lut) hipFree(b->d_lut);
  if (b->d_rankLo) hipFree(b->d_rankLo);
  if (b->d_rankHi) hipFree(b->d_rankHi);
  if (b->d_rankLo16) hipFree(b->d_rankLo16);
  if (b->d_binom) hipFree(b->d_binom);
  delete b;
  return DH_OK;
}

extern "C" int64_t dh_basis_dimension(const dh_basis* b) { return b ? b->dim : -1; }

__global__ void widen_words_kernel(const uint32_t* __restrict__ w, uint64_t* __restrict__ out, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (uint64_t)w[i];
}

extern "C" int dh_fermion_sphere_from_states(dh_context* c, dh_basis** out, int N, int totalLz, int lzMax,
                                             const uint64_t* states, int64_t dim) {
  if (!c || !out || !states) return DH_E_ARG;
  *out = nullptr;
  const int64_t full = dh_fermion_sphere_dimension(N, totalLz, lzMax);
  if (dim != full)
    return dh_fail(c, DH_E_ARG, "the states are not the full sector: " + std::to_string(dim) + " words for a sector of " +
                                    std::to_string(full) + " states");
  dh_basis* b = nullptr;
  int rc = dh_fermion_sphere_create(c, &b, N, totalLz, lzMax);
  if (rc) return rc;
  std::vector<uint64_t> mine((size_t)dim);
  if ((rc = dh_basis_get_states(b, mine.data()))) {
    dh_basis_destroy(b);
    return rc;
  }
  for (int64_t i = 0; i < dim; ++i)
    if (mine[(size_t)i] != states[i]) {
      dh_basis_destroy(b);
      return dh_fail(c, DH_E_ARG, "state " + std::to_string(i) + " differs from the sector's state in the reference's order");
    }
  *out = b;
  return DH_OK;
}

extern "C" int dh_basis_get_states(dh_basis* b, uint64_t* states) {
  if (!b || !states) return DH_E_ARG;
  dh_context* c = b->ctx;
  if (b->kind == DH_BASIS_FERMION_SPHERE) {
    DH_HIP(c, hipMemcpyAsync(states, b->d_states, sizeof(uint64_t) * b->dim, hipMemcpyDeviceToHost, c->stream));
  } else {
    uint64_t* tmp;
    int rc = dh_alloc(c, (void**)&tmp, sizeof(uint64_t) * b->dim);
    if (rc) return rc;
    hipLaunchKernelGGL(widen_words_kernel, dim3((unsigned)((b->dim + 255) / 256)), dim3(256), 0, c->stream,
                       b->d_words, tmp, b->dim);
    DH_HIP(c, hipMemcpyAsync(states, tmp, sizeof(uint64_t) * b->dim, hipMemcpyDeviceToHost, c->stream));
    DH_HIP(c, hipStreamSynchronize(c->stream));
    hipFree(tmp);
  }
  DH_HIP(c, hipStreamSynchronize(c->stream));
  return DH_OK;
}
