// Internal structures shared by the libdhgpu translation units (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <type_traits>
#include <unordered_set>
#include <vector>

#include "../../include/dhgpu.h"

// A context is reference counted.  dh_create hands its caller one reference, and every basis,
// Hamiltonian, sector set and dh_malloc buffer of the context holds one more, so the context (its
// stream, scratch and staging memory) stays valid until the last of them is released, whatever
// order their owners release them in.  dh_destroy gives up the caller's reference.
struct dh_context {
  std::atomic<int> refs{1};
  bool closed = false;  // dh_destroy was called
  std::mutex allocMutex;
  std::unordered_set<void*> allocs;  // live dh_malloc buffers
  int device = 0;
  hipStream_t stream = nullptr;
  bool ownStream = false;
  std::string lastError;
  int nbrCU = 0;
  char arch[64] = {0};
  // reduction scratch (device) and pinned staging (host), grown on demand
  double* d_scratch = nullptr;
  size_t scratchBytes = 0;
  double* h_pinned = nullptr;
  size_t pinnedBytes = 0;
  // interleaved staging block of dh_hv_vectors (device), grown on demand
  double* d_pack = nullptr;
  size_t packBytes = 0;
  // device memory in use (all of the device's, sampled after each allocation of this context) at
  // its highest since creation or the last dh_memory_peak reset
  size_t usedPeak = 0;
  // ring of pinned slots for host-to-device uploads that do not wait: a slot is reused once the
  // event recorded after its copy has completed
  static constexpr int kStageSlots = 16;
  struct StageSlot {
    void* host = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool pending = false;
  } stage[kStageSlots];
  int stageNext = 0;
};

void dh_ctx_retain(dh_context* c);
void dh_ctx_release(dh_context* c);

// Counted reference to a context, held by every object that uses one.  Assigning a context takes
// a reference and destroying the holder releases it, so every path that deletes the object
// (its destroy call or a failed creation) releases the context too.
class dh_ctx_ref {
 public:
  dh_ctx_ref() = default;
  dh_ctx_ref(const dh_ctx_ref&) = delete;
  dh_ctx_ref& operator=(const dh_ctx_ref&) = delete;
  ~dh_ctx_ref() {
    if (p) dh_ctx_release(p);
  }
  dh_ctx_ref& operator=(dh_context* c) {
    if (c) dh_ctx_retain(c);
    if (p) dh_ctx_release(p);
    p = c;
    return *this;
  }
  operator dh_context*() const { return p; }
  dh_context* operator->() const { return p; }

 private:
  dh_context* p = nullptr;
};

enum dh_basis_kind { DH_BASIS_FERMION_SPHERE = 1, DH_BASIS_SPIN_CHAIN = 2 };

struct dh_basis {
  dh_ctx_ref ctx;
  int kind = 0;
  int64_t dim = 0;
  // fermions on the sphere
  int nbrFermions = 0, lzMax = 0, totalLz = 0, nbrOrbitals = 0, orbitalSum = 0;
  uint64_t* d_states = nullptr;  // [dim] descending uint64
  // spin chain
  int chainLength = 0, nbrUp = 0, sz = 0;
  uint32_t* d_words = nullptr;  // [dim] reference 32-bit words (bits >= L set)
  // LUT for the binary-search lookup: lut[v] = number of states whose top lutBits bits
  // (of the ordering key) are >= v, v in [0, 2^lutBits]
  int lutBits = 0, lutShift = 0;
  int* d_lut = nullptr;
  // combinatorial rank tables (fermions): index = dim - 1 - (lo[s & (2^rankBits-1)] + hi[s >> rankBits]);
  // d_rankLo16: the same low table in 16 bits (null when a value does not fit), for LDS staging
  int rankBits = 0;
  int64_t rankLoSize = 0, rankHiSize = 0;
  uint32_t* d_rankLo = nullptr;
  uint32_t* d_rankHi = nullptr;
  uint16_t* d_rankLo16 = nullptr;
  // binomial table C[n][k], n,k <= 32 (spin rank)
  uint32_t* d_binom = nullptr;
};

enum dh_ham_kind { DH_HAM_SPHERE = 1, DH_HAM_SPIN = 2, DH_HAM_SPHERE_NBODY = 3 };

struct dh_hamiltonian {
  dh_ctx_ref ctx;
  dh_basis* basis = nullptr;
  int kind = 0;
  double shift = 0.0;
  // sphere two-body factors
  int nbrFactors = 0, nbrPairs = 0;
  double* d_factors = nullptr;
  int* d_pairOffset = nullptr;  // [nbrPairs + 1]
  // the same factors packed symmetric per pair sum (what the product kernels stage in LDS)
  int nbrPackedFactors = 0, nbrPairSums = 0;
  double* d_packedFactors = nullptr;
  int* d_sumBase = nullptr;  // [nbrPairSums]
  // adjoint channel split
  int nbrChannels = 0;
  double* d_cg = nullptr;  // [nbrPairs * nbrChannels]
  int64_t* d_adjBinom = nullptr;  // adjoint: C(j, m), j <= nbrOrbitals, m <= nbrFermions - 2
  std::vector<double> channelWeights;
  // n-body term with explicit per-sum factor matrices: tuples as orbital bit masks grouped by
  // orbital sum, the position of each tuple in its sum (by combinatorial rank of the tuple)
  int nbody = 0, nbrTuples = 0, nbrTupleSums = 0;
  int64_t tuplePosSize = 0, nbrNBodyFactors = 0;
  uint64_t* d_tupleMask = nullptr;
  int* d_tuplePos = nullptr;         // [C(nbrOrbitals, nbody)], -1 if the tuple is not listed
  int* d_sumTupleOff = nullptr;      // [nbrTupleSums + 1]
  int64_t* d_sumFactorOff = nullptr;  // [nbrTupleSums + 1]
  double* d_nbodyFactors = nullptr;
  // stored matrix elements (dh_hamiltonian_precompute): CSR of the operator without the shift
  int64_t csrNnz = 0;
  int64_t* d_csrRow = nullptr;  // [dim + 1]
  int32_t* d_csrCol = nullptr;  // [csrNnz]
  double* d_csrVal = nullptr;   // [csrNnz]
  // spin chain couplings
  double J[32] = {0}, Jz[32] = {0};
  int periodic = 0;
  // launch configuration
  int lanesPerRow = 1, blockSize = 256, indexMode = DH_INDEX_RANK, ldsTables = 1;
  // sphere: pick lanes per row and workgroup size per product from k (until set_launch is called)
  bool autoLaunch = false;
};

enum dh_sector_kind { DH_SECTORS_SPHERE = 1, DH_SECTORS_MOMENTUM = 2, DH_SECTORS_TORUS_TRANSLATIONS = 3 };

// A set of symmetry sectors of one problem with concatenated bases (sector s owns global
// rows [offsets[s], offsets[s+1])).  Sphere sets are real (Lz sectors); momentum sets are
// complex ((kx, ky) sectors, vectors stored as interleaved (re, im) pairs).
struct dh_sector_set {
  dh_ctx_ref ctx;
  int kind = DH_SECTORS_SPHERE;
  int nbrFermions = 0, lzMax = 0, nbrOrbitals = 0, nbrSectors = 0;
  std::vector<int> totalLz, orbitalSum;  // sphere: 2 Lz and orbital-index sum per sector
  int nbrSiteX = 0, nbrSiteY = 0;        // momentum: lattice and (kx, ky) per sector
  std::vector<int> kx, ky;
  std::vector<int64_t> offsets;  // nbrSectors + 1
  int64_t totalDim = 0;
  uint64_t* d_states = nullptr;
  int rankBits = 0;
  int64_t loSize = 0, hiSize = 0;  // hi table of sector s starts at s * hiSize
  uint32_t* d_rankLo = nullptr;
  uint32_t* d_rankHi = nullptr;
  uint16_t* d_rankLo16 = nullptr;  // the low table in 16 bits (null when a value does not fit)
  int64_t* d_offsets = nullptr;
  // sphere factors (per-pair layout, and packed symmetric per pair sum)
  double* d_factors = nullptr;
  int* d_pairOffset = nullptr;
  int nbrFactors = 0, nbrPairs = 0;
  int nbrPackedFactors = 0, nbrPairSums = 0;
  double* d_packedFactors = nullptr;
  int* d_sumBase = nullptr;
  // momentum factors: per momentum-sum sector j its pair list and complex factor matrix
  int nbrSums = 0;
  int* d_pairInfo = nullptr;     // [pair id b(b-1)/2 + a] -> (sum << 16) | position
  int* d_sumPairOff = nullptr;   // [nbrSums + 1] into d_sumPairs
  int* d_sumPairs = nullptr;     // (a, b) packed as a | (b << 8)
  int64_t* d_sumFactorOff = nullptr;  // [nbrSums + 1] complex entries
  double* d_cfactors = nullptr;  // interleaved (re, im)
  int64_t nbrComplexFactors = 0;
  // free-pair tables of a momentum set, per momentum sum j (see lattice_free_pairs): masks
  // [3 j] = positions whose ky stays in its block / wraps / has the smaller orbital of its pair,
  // shifts [2 j] = (kx rotation in bits, ky rotation), and per orbital m the pair (i2 << 8) | m2
  uint64_t* d_freeMask = nullptr;
  int* d_freeShift = nullptr;
  int* d_partnerPos = nullptr;
  double shift = 0.0;
  // torus with magnetic translations: representatives ascending per sector, their orbit sizes and
  // reordering-sign bits, a top-bits lookup per sector, and the phases exp(2 pi i Kx t / modulo)
  uint8_t* d_orbit = nullptr;
  uint64_t* d_reorder = nullptr;
  int32_t* d_lut = nullptr;    // per sector (1 << lutBits) + 1 first rows of each top-bits prefix
  double* d_phase = nullptr;   // per sector `modulo` (cos, sin) pairs
  int lutBits = 0, modulo = 1, stateShift = 0;
  int lanesPerRow = 1, blockSize = 256;
  int rowMode = DH_INDEX_RANK;  // sphere sets: DH_INDEX_RANK or DH_INDEX_RANK_FREE row routine
  // segmented BLAS launch shape: persistent axpy workgroups per CU, dots rows per tile (0: auto)
  int axpyGroupsPerCU = 16;
  int64_t dotsRowsPerTile = 0;
  // per-call work descriptor: [active sector ids (int64)] [tile prefix (int64)]
  int64_t* d_work = nullptr;
  size_t workCap = 0;
  double* d_coef = nullptr;  // per-sector coefficient matrices for axpy / rmul
  size_t coefCap = 0;
};

// work descriptor of the active sectors (dh_multisector.hip)
int dh_sectors_upload_work(dh_sector_set* S, const unsigned char* active, int64_t rowsPerTile, int* nActive,
                           int64_t* totalTiles);
// complex product over a momentum set and its matrix elements per sector (dh_lattice.hip)
int dh_lattice_sectors_hv(dh_sector_set* S, const double* x, double* y, int k, int add, const unsigned char* active);
int dh_lattice_count_elements(dh_sector_set* S, int64_t* counts);
// complex product over a torus set with magnetic translations (dh_torus.hip)
int dh_torus_translation_sectors_hv(dh_sector_set* S, const double* x, double* y, int k, int add,
                                    const unsigned char* active);
// product from stored matrix elements and their release (dh_aux.hip)
int dh_csr_hv(dh_hamiltonian* h, const double* x, double* y, int k, int add, int64_t first, int64_t count);
void dh_csr_free(dh_hamiltonian* h);
// n-body product and element count (dh_nbody.hip)
int dh_nbody_hv(dh_hamiltonian* h, const double* x, double* y, int k, int add, int64_t first, int64_t count);
int dh_nbody_count_elements(dh_hamiltonian* h, int64_t* nnz);
void dh_nbody_free(dh_hamiltonian* h);

// error helpers
int dh_fail(dh_context* ctx, int code, const std::string& msg);
#define DH_HIP(ctx, call)                                                                 \
  do {                                                                                    \
    hipError_t _e = (call);                                                               \
    if (_e != hipSuccess)                                                                 \
      return dh_fail((ctx), DH_E_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)

// device memory owned by the library's own objects (not counted as dh_malloc buffers)
int dh_alloc(dh_context* ctx, void** d_ptr, size_t bytes);
// scratch management (core)
int dh_scratch(dh_context* ctx, size_t bytes, double** d_ptr);
int dh_pinned(dh_context* ctx, size_t bytes, double** h_ptr);
// copy `bytes` of host memory to device memory in stream order without waiting for it: the host
// data is staged in a pinned slot first, so the caller may reuse it as soon as this returns
int dh_upload_async(dh_context* ctx, void* d_dst, const void* h_src, size_t bytes);

// Lanes that share one row in the block dot products: the K x K products of a row are split over
// QP x QQ lanes (rows of the basis block x columns of W), so a lane holds at most 16 accumulators
// per basis block.  One lane per row would hold K*K: 64 doubles at K = 8 (166 VGPRs, three
// wavefronts per SIMD) and 256 at K = 16 (spilled to scratch, one wavefront per SIMD).
template <int K>
struct DotsLanes {
  static constexpr int QP = K >= 16 ? 4 : (K >= 8 ? 2 : 1);
  static constexpr int QQ = QP;
  static constexpr int QS = QP * QQ;
};

// Krylov basis rows are read once per pass and never again before the next pass, so at k = 1 they
// are loaded with the nontemporal hint (no allocation in the caches, which keep the rows of W that
// several workgroups share): 6.2-6.5 against 5.3-5.8 TB/s for the k = 1 dot products on MI355X.
// Blocks of k > 1 keep ordinary loads: lanes that split a row's products read the same basis
// elements, and the block solves ran 7-15 % slower with the hint.
typedef double dh_double2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double dh_stream_load(const double* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ dh_double2 dh_stream_load2(const double* p) {
  return __builtin_nontemporal_load(reinterpret_cast<const dh_double2*>(p));
}

// k = 1 dot products over rows [rowBegin, rowEnd) two rows per lane with 16-byte loads (V and W
// 16-byte aligned, blockStride even): the aligned row pairs inside the range, then an odd first
// row and an odd last row on lanes 0 and 1.  acc[g] += sum of V_{b0+g}[row] W[row].
template <int G>
__device__ __forceinline__ void dh_dots_rows_pairs(const double* __restrict__ V, int64_t blockStride, int nbrBlocks,
                                                   int b0, const double* __restrict__ W, int64_t rowBegin,
                                                   int64_t rowEnd, double* acc) {
  const int64_t pb = (rowBegin + 1) >> 1, pe = rowEnd >> 1;
  const int ng = min(G, nbrBlocks - b0);
  const double* vb = V + (int64_t)b0 * blockStride;
  if (ng == G) {
    for (int64_t q = pb + threadIdx.x; q < pe; q += blockDim.x) {
      const dh_double2 w = *reinterpret_cast<const dh_double2*>(W + 2 * q);
      dh_double2 v[G];
#pragma unroll
      for (int g = 0; g < G; ++g) v[g] = dh_stream_load2(vb + g * blockStride + 2 * q);
#pragma unroll
      for (int g = 0; g < G; ++g) acc[g] += v[g].x * w.x + v[g].y * w.y;
    }
  } else {
    for (int64_t q = pb + threadIdx.x; q < pe; q += blockDim.x) {
      const dh_double2 w = *reinterpret_cast<const dh_double2*>(W + 2 * q);
#pragma unroll
      for (int g = 0; g < G; ++g)
        if (g < ng) {
          const dh_double2 v = dh_stream_load2(vb + g * blockStride + 2 * q);
          acc[g] += v.x * w.x + v.y * w.y;
        }
    }
  }
  int64_t single = -1;
  if (threadIdx.x == 0 && (rowBegin & 1) && rowBegin < rowEnd) single = rowBegin;
  if (threadIdx.x == 1 && (rowEnd & 1) && rowEnd - 1 >= rowBegin) single = rowEnd - 1;
  if (single >= 0) {
    const double w = W[single];
#pragma unroll
    for (int g = 0; g < G; ++g)
      if (g < ng) acc[g] += vb[g * blockStride + single] * w;
  }
}

// Sums over rows [rowBegin, rowEnd) of V_b[row][p] W[row][q] for the basis blocks b0 .. b0+G-1
// (block b at V + b * blockStride, rows of K doubles), written b-major then p then q to out[],
// for the blocks below nbrBlocks.  Every thread of the workgroup must call it (one barrier);
// red is LDS for (blockDim.x / 64) * G * K * K doubles.  pairs (k = 1 only): V and W 16-byte
// aligned and blockStride even, so rows are read in pairs.
template <int K, int G>
__device__ __forceinline__ void dh_dots_tile(const double* __restrict__ V, int64_t blockStride, int nbrBlocks,
                                             int b0, const double* __restrict__ W, int64_t rowBegin,
                                             int64_t rowEnd, double* red, double* __restrict__ out,
                                             bool pairs = false) {
  constexpr int QP = DotsLanes<K>::QP, QQ = DotsLanes<K>::QQ, QS = DotsLanes<K>::QS;
  constexpr int KP = K / QP, KQ = K / QQ, NA = G * KP * KQ, KK = K * K;
  const int sub = threadIdx.x % QS, sp = sub / QQ, sq = sub % QQ;
  double acc[NA];
#pragma unroll
  for (int i = 0; i < NA; ++i) acc[i] = 0.0;
  if (K == 1 && pairs) {
    dh_dots_rows_pairs<G>(V, blockStride, nbrBlocks, b0, W, rowBegin, rowEnd, acc);
  } else {
    for (int64_t row = rowBegin + threadIdx.x / QS; row < rowEnd; row += blockDim.x / QS) {
      double w[KQ];
#pragma unroll
      for (int qq = 0; qq < KQ; ++qq) w[qq] = W[row * K + sq * KQ + qq];
#pragma unroll
      for (int g = 0; g < G; ++g)
        if (b0 + g < nbrBlocks) {
          const double* vb = V + (int64_t)(b0 + g) * blockStride + row * K + sp * KP;
#pragma unroll
          for (int pp = 0; pp < KP; ++pp) {
            const double v = K == 1 ? dh_stream_load(vb + pp) : vb[pp];
#pragma unroll
            for (int qq = 0; qq < KQ; ++qq) acc[(g * KP + pp) * KQ + qq] += v * w[qq];
          }
        }
    }
  }
  // sum over the lanes of a wavefront with the same share (lane mod QS), then over wavefronts
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NA; ++i) {
    double v = acc[i];
    for (int o = 32; o >= QS; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane < QS) red[(wave * QS + lane) * NA + i] = v;
  }
  __syncthreads();
  const int nw = (int)(blockDim.x >> 6);
  for (int idx = threadIdx.x; idx < G * KK; idx += blockDim.x) {
    const int g = idx / KK, p = (idx % KK) / K, q = idx % K;
    if (b0 + g >= nbrBlocks) continue;
    const int s = (p / KP) * QQ + q / KQ, i = (g * KP + p % KP) * KQ + q % KQ;
    double t = 0.0;
    for (int w2 = 0; w2 < nw; ++w2) t += red[(w2 * QS + s) * NA + i];
    out[idx] = t;
  }
}

// basis blocks per workgroup pass of the dot products (sharing one read of W): 8 at k = 1 (faster
// than 16 on MI355X: 5.3 against 4.9 TB/s before the row pairs), else 16 / k^2, at least one
template <int K>
struct DotsGroup {
  static constexpr int value = K == 1 ? 8 : ((16 / (K * K)) > 0 ? (16 / (K * K)) : 1);
};

// whether the k = 1 dot products may read rows in pairs
inline bool dh_pairs_aligned(const double* V, int64_t blockStride, const double* W) {
  return blockStride % 2 == 0 && reinterpret_cast<uintptr_t>(V) % 16 == 0 && reinterpret_cast<uintptr_t>(W) % 16 == 0;
}

// Workgroups of `kern` that one CU holds at blockSize threads and `lds` bytes of dynamic LDS (at
// least one).  Asked of the runtime once per kernel, shape and device and then remembered: the
// query takes about 11 us, as long as a small sector's whole product.
template <typename Kern>
inline int dh_blocks_per_cu(const dh_context* c, Kern kern, int blockSize, size_t lds) {
  static std::mutex mutex;
  static std::map<std::tuple<const void*, int, size_t, int>, int> cache;
  const auto key = std::make_tuple(reinterpret_cast<const void*>(kern), blockSize, lds, c->device);
  {
    std::lock_guard<std::mutex> lock(mutex);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  int perCU = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kern, blockSize, lds) != hipSuccess || perCU < 1) perCU = 1;
  std::lock_guard<std::mutex> lock(mutex);
  cache[key] = perCU;
  return perCU;
}

// XCD-aware block remap: MI355X hands consecutive workgroups to its 8 XCDs round robin;
// remapping gives every XCD one contiguous range of blocks so neighbouring rows (whose
// gather targets are close) share that XCD's L2.
__device__ __forceinline__ unsigned dh_xcd_block(unsigned b, unsigned n) {
  const unsigned nx = 8;
  if (n < nx) return b;
  unsigned per = n / nx, rem = n % nx;
  unsigned x = b % nx, i = b / nx;
  // XCD x owns blocks [x*per + min(x,rem), ...) of length per + (x < rem)
  unsigned start = x * per + (x < rem ? x : rem);
  return start + i;
}

// Two-body factors on the sphere are symmetric, F(a -> b) = F(b -> a), between the pairs of one
// orbital sum S = n1 + n2.  Pair (n1 > n2) of sum S has position n1 - lo(S), lo(S) = S/2 + 1, and
// the factor between positions i and j (i <= j) is packed at base[S] + j (j + 1) / 2 + i.  This
// halves the table the product kernels stage in LDS.  Input: the per-pair layout (pair
// a = n1 (n1 - 1) / 2 + n2, created m3 from lo(S): factors[pairOffset[a] + m3 - lo(S)]).
inline void dh_pack_sphere_factors(int nbrOrbitals, const double* factors, const int* pairOffset,
                                   std::vector<double>& packed, std::vector<int>& base) {
  const int nbrSums = 2 * nbrOrbitals - 1;
  packed.clear();
  base.assign(nbrSums, 0);
  for (int S = 0; S < nbrSums; ++S) {
    base[S] = (int)packed.size();
    const int lo = S / 2 + 1, hi = std::min(S, nbrOrbitals - 1);
    for (int j = 0; lo + j <= hi; ++j)
      for (int i = 0; i <= j; ++i) {
        const int n1 = lo + i, n2 = S - n1;
        packed.push_back(factors[pairOffset[n1 * (n1 - 1) / 2 + n2] + j]);
      }
  }
}

// packs, and refuses factor tables that are not symmetric (the products read one triangle)
inline int dh_pack_factors_checked(dh_context* c, int nbrOrbitals, const double* factors, const int* pairOffset,
                                   std::vector<double>& packed, std::vector<int>& base) {
  dh_pack_sphere_factors(nbrOrbitals, factors, pairOffset, packed, base);
  double scale = 0.0, asym = 0.0;
  for (int S = 0; S < 2 * nbrOrbitals - 1; ++S) {
    const int lo = S / 2 + 1, hi = std::min(S, nbrOrbitals - 1);
    for (int i = 0; lo + i <= hi; ++i)
      for (int j = 0; lo + j <= hi; ++j) {
        const int n1 = lo + i, n2 = S - n1;
        const double f = factors[pairOffset[n1 * (n1 - 1) / 2 + n2] + j];
        const int m3 = lo + j, m4 = S - m3;
        const double g = factors[pairOffset[m3 * (m3 - 1) / 2 + m4] + i];
        scale = std::max(scale, std::fabs(f));
        asym = std::max(asym, std::fabs(f - g));
      }
  }
  if (asym > 1e-12 * scale) return dh_fail(c, DH_E_ARG, "factors are not symmetric between annihilated and created pairs");
  return DH_OK;
}

__host__ __device__ __forceinline__ int dh_packed_index(int i, int j) {
  return i <= j ? ((j * (j + 1)) >> 1) + i : ((i * (i + 1)) >> 1) + j;
}

// binomial C(n, k) for the host, 64-bit
inline uint64_t dh_binom_host(int n, int k) {
  if (k < 0 || k > n) return 0;
  uint64_t r = 1;
  for (int i = 1; i <= k; ++i) r = r * (uint64_t)(n - k + i) / (uint64_t)i;
  return r;
}

// Low-table width of the combinatorial rank (index = dim - 1 - lo[low bits] - hi[high bits]).  The
// low table counts patterns of the low orbitals with the same particle number and momentum, so its
// values stay below C(bits, bits / 2) and fit 16 bits: at 14 bits it is 32 KiB and is staged in LDS
// by the product kernels.  Up to 28 orbitals the split is even; from 29 to 34 orbitals the low part
// stays at 14 bits and the high table grows to at most 2^20 entries (4 MiB, of which a sector
// touches a small clustered part: 1.6 MB for N=12 2S=33 Lz=0); beyond 34 the split is even again.
inline int dh_rank_split(int nbrOrbitals) {
  return (nbrOrbitals > 28 && nbrOrbitals <= 34) ? 14 : (nbrOrbitals + 1) / 2;
}

// f(i) for i in [0, n) on the host's hardware threads (table construction on the host)
template <class F>
inline void dh_parallel_for(int n, F f) {
  const int nt = std::max(1, std::min<int>(n, (int)std::thread::hardware_concurrency()));
  if (nt <= 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<int> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&] {
      for (int i = next++; i < n; i = next++) f(i);
    });
  for (auto& th : pool) th.join();
}

// 16-bit copy of a low rank table, or an empty vector when the table is longer than 2^15 entries or a
// value does not fit
inline std::vector<uint16_t> dh_rank_lo16(const std::vector<uint32_t>& lo) {
  std::vector<uint16_t> out;
  if (lo.size() > ((size_t)1 << 15)) return out;
  out.reserve(lo.size());
  for (uint32_t v : lo) {
    if (v > 0xffffu) return {};
    out.push_back((uint16_t)v);
  }
  return out;
}
