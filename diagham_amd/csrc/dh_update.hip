// One pass over the Krylov basis for the middle of the two-round block Gram-Schmidt:
//   W <- W M + sum_b V_b C_b          (M optional, k x k per segment)
//   D_b = V_b^T W, G = W^T W          (over the updated W; D optional)
// for blocks of KR <= 2 doubles per row (one real or one complex vector).  Separate axpy and dots
// kernels read the basis twice; here a workgroup stages the basis rows of a small row tile in LDS,
// forms the new W rows from it, then takes the dots from the same LDS copy, so the basis is read
// from HBM once.  A workgroup walks the row tiles of one chunk (rows of one segment) and keeps its
// dot sums in registers; one partial per chunk is reduced per segment afterwards.
#include <algorithm>
#include <cstring>
#include <vector>

#include "dh_internal.h"

namespace {
constexpr int kThreads = 256;
constexpr int kTileDoubles = 5120;  // 40 KiB of LDS for the basis tile
constexpr int kMaxOutPerThread = 8;  // dot outputs (b, p, q) per thread
}  // namespace

// chunk c: rows [start[c], end[c]) of segment seg[c] (index among the active segments)
template <int KR>
__global__ void __launch_bounds__(kThreads) block_update_kernel(const double* __restrict__ V, int nb,
                                                                int64_t blockStride, double* __restrict__ W,
                                                                const int64_t* __restrict__ chunkStart,
                                                                const int64_t* __restrict__ chunkEnd,
                                                                const int* __restrict__ chunkSeg,
                                                                const double* __restrict__ C,
                                                                const double* __restrict__ M, int wantDots, int R,
                                                                double* __restrict__ partial) {
  constexpr int KK = KR * KR;
  __shared__ double sV[kTileDoubles];
  __shared__ double sW[kThreads];          // new W rows of the tile (R * KR <= 256)
  __shared__ double sPart[kThreads];       // phase-1 partial sums
  __shared__ double sM[KK];
  extern __shared__ double sC[];           // nb * KK coefficients of this chunk's segment
  const int chunk = blockIdx.x;
  const int64_t start = chunkStart[chunk], end = chunkEnd[chunk];
  const int seg = chunkSeg[chunk];
  const int items = R * KR;       // W values per tile
  const int pitch = R * KR + 1;   // LDS row pitch of one basis block (odd: no bank aliasing)
  const int groups = kThreads / items;
  for (int i = threadIdx.x; i < nb * KK; i += kThreads) sC[i] = C[(int64_t)seg * nb * KK + i];
  if (threadIdx.x < KK) sM[threadIdx.x] = M ? M[(int64_t)seg * KK + threadIdx.x] : (threadIdx.x % (KR + 1) == 0 ? 1.0 : 0.0);
  const int nOut = (wantDots ? nb * KK : 0) + KK;
  double acc[kMaxOutPerThread];
#pragma unroll
  for (int o = 0; o < kMaxOutPerThread; ++o) acc[o] = 0.0;
  for (int64_t r0 = start; r0 < end; r0 += R) {
    const int rows = (int)min<int64_t>(R, end - r0);
    __syncthreads();  // previous tile's readers are done (and sC / sM are loaded)
    // phase 0: basis rows of the tile into LDS, coalesced per block
    for (int i = threadIdx.x; i < nb * items; i += kThreads) {
      const int b = i / items, e = i - b * items;
      sV[b * pitch + e] = (e < rows * KR) ? V[(int64_t)b * blockStride + r0 * KR + e] : 0.0;
    }
    __syncthreads();
    // phase 1: new W values, the block sum split over `groups` thread groups
    {
      const int item = threadIdx.x % items, g = threadIdx.x / items;
      const int r = item / KR, q = item - r * KR;
      double s = 0.0;
      if (g < groups)
        for (int b = g; b < nb; b += groups) {
#pragma unroll
          for (int p = 0; p < KR; ++p) s += sV[b * pitch + r * KR + p] * sC[b * KK + p * KR + q];
        }
      sPart[threadIdx.x] = s;
    }
    __syncthreads();
    if (threadIdx.x < items) {
      const int r = threadIdx.x / KR, q = threadIdx.x - r * KR;
      double s = 0.0;
      if (r < rows) {
        const double* w = W + (r0 + r) * KR;
#pragma unroll
        for (int p = 0; p < KR; ++p) s += w[p] * sM[p * KR + q];
        for (int g = 0; g < groups; ++g) s += sPart[g * items + threadIdx.x];
      }
      sW[threadIdx.x] = s;
    }
    __syncthreads();
    if (threadIdx.x < rows * KR) W[r0 * KR + threadIdx.x] = sW[threadIdx.x];
    // phase 2: dot outputs o = (b, p, q) of V_b^T W, then (nb, p, q) of W^T W
#pragma unroll
    for (int k = 0; k < kMaxOutPerThread; ++k) {
      const int o = threadIdx.x + k * kThreads;
      if (o < nOut) {
        const int b = o / KK, pq = o - b * KK, p = pq / KR, q = pq - p * KR;
        const bool gram = !wantDots || b == nb;
        double s = 0.0;
        for (int r = 0; r < rows; ++r)
          s += (gram ? sW[r * KR + p] : sV[b * pitch + r * KR + p]) * sW[r * KR + q];
        acc[k] += s;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kMaxOutPerThread; ++k) {
    const int o = threadIdx.x + k * kThreads;
    if (o < nOut) partial[(int64_t)chunk * nOut + o] = acc[k];
  }
}

// per segment j: out[j][o] = sum of partial[chunk][o] over its chunks [first[j], first[j+1])
__global__ void block_update_reduce_kernel(const double* __restrict__ partial, const int* __restrict__ first,
                                           int nOut, double* __restrict__ out) {
  const int j = blockIdx.y;
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nOut) return;
  double s = 0.0;
  for (int c = first[j]; c < first[j + 1]; ++c) s += partial[(int64_t)c * nOut + o];
  out[(int64_t)j * nOut + o] = s;
}

// rows per tile for nb blocks of KR doubles per row, 0 when the fused pass does not apply
static int update_tile_rows(int nb, int kr, bool wantDots) {
  if (kr != 1 && kr != 2) return 0;
  const int nOut = (wantDots ? nb * kr * kr : 0) + kr * kr;
  if (nOut > kThreads * kMaxOutPerThread) return 0;
  int R = kThreads / kr;
  while (R > 8 && nb * (R * kr + 1) > kTileDoubles) R >>= 1;
  if (nb * (R * kr + 1) > kTileDoubles) return 0;
  return R;
}

// segments: row ranges [segStart[j], segEnd[j]) of the active segments; C: nb * KK per segment,
// M: KK per segment or null; dots (nb * KK per segment) may be null; gram: KK per segment
int dh_update_segments(dh_context* c, const double* V, int nb, int64_t blockStride, double* W, int kr,
                       const std::vector<int64_t>& segStart, const std::vector<int64_t>& segEnd, const double* h_C,
                       const double* h_M, double* h_dots, double* h_gram) {
  const bool wantDots = h_dots != nullptr;
  const int R = update_tile_rows(nb, kr, wantDots);
  if (R == 0) return DH_E_RANGE;  // the caller takes the separate kernels
  const int na = (int)segStart.size();
  const int KK = kr * kr;
  const int nOut = (wantDots ? nb * KK : 0) + KK;
  // chunks of up to 8192 rows (whole tiles) inside each segment
  const int64_t chunkRows = std::max<int64_t>(R, (8192 / R) * R);
  std::vector<int64_t> cs, ce;
  std::vector<int> cseg, first(na + 1, 0);
  for (int j = 0; j < na; ++j) {
    first[j] = (int)cs.size();
    for (int64_t r = segStart[j]; r < segEnd[j]; r += chunkRows) {
      cs.push_back(r);
      ce.push_back(std::min(segEnd[j], r + chunkRows));
      cseg.push_back(j);
    }
  }
  first[na] = (int)cs.size();
  const int nChunks = (int)cs.size();
  if (nChunks == 0) {
    if (h_dots) std::memset(h_dots, 0, sizeof(double) * (size_t)na * nb * KK);
    std::memset(h_gram, 0, sizeof(double) * (size_t)na * KK);
    return DH_OK;
  }
  // one scratch area: descriptors, coefficients, partials, sums
  const size_t descDoubles = (size_t)nChunks * 2 + ((size_t)nChunks + na + 1 + 1) / 2 + 1;
  const size_t coefDoubles = (size_t)na * nb * KK + (h_M ? (size_t)na * KK : 0);
  const size_t partDoubles = (size_t)nChunks * nOut, outDoubles = (size_t)na * nOut;
  double* scratch;
  int rc = dh_scratch(c, sizeof(double) * (descDoubles + coefDoubles + partDoubles + outDoubles), &scratch);
  if (rc) return rc;
  int64_t* d_cs = reinterpret_cast<int64_t*>(scratch);
  int64_t* d_ce = d_cs + nChunks;
  int* d_cseg = reinterpret_cast<int*>(d_ce + nChunks);
  int* d_first = d_cseg + nChunks;
  double* d_C = scratch + descDoubles;
  double* d_M = h_M ? d_C + (size_t)na * nb * KK : nullptr;
  double* d_part = d_C + coefDoubles;
  double* d_out = d_part + partDoubles;
  std::vector<int> ints(cseg);
  ints.insert(ints.end(), first.begin(), first.end());
  if ((rc = dh_upload_async(c, d_cs, cs.data(), sizeof(int64_t) * nChunks)) ||
      (rc = dh_upload_async(c, d_ce, ce.data(), sizeof(int64_t) * nChunks)) ||
      (rc = dh_upload_async(c, d_cseg, ints.data(), sizeof(int) * ints.size())) ||
      (rc = dh_upload_async(c, d_C, h_C, sizeof(double) * (size_t)na * nb * KK)) ||
      (h_M && (rc = dh_upload_async(c, d_M, h_M, sizeof(double) * (size_t)na * KK))))
    return rc;
  const size_t lds = sizeof(double) * (size_t)nb * KK;
  if (kr == 1)
    hipLaunchKernelGGL(block_update_kernel<1>, dim3(nChunks), dim3(kThreads), lds, c->stream, V, nb, blockStride, W,
                       d_cs, d_ce, d_cseg, d_C, d_M, (int)wantDots, R, d_part);
  else
    hipLaunchKernelGGL(block_update_kernel<2>, dim3(nChunks), dim3(kThreads), lds, c->stream, V, nb, blockStride, W,
                       d_cs, d_ce, d_cseg, d_C, d_M, (int)wantDots, R, d_part);
  hipLaunchKernelGGL(block_update_reduce_kernel, dim3((nOut + 255) / 256, na), dim3(256), 0, c->stream, d_part,
                     d_first, nOut, d_out);
  DH_HIP(c, hipGetLastError());
  double* pin;
  if ((rc = dh_pinned(c, sizeof(double) * outDoubles, &pin))) return rc;
  DH_HIP(c, hipMemcpyAsync(pin, d_out, sizeof(double) * outDoubles, hipMemcpyDeviceToHost, c->stream));
  DH_HIP(c, hipStreamSynchronize(c->stream));
  for (int j = 0; j < na; ++j) {
    if (h_dots) std::memcpy(h_dots + (size_t)j * nb * KK, pin + (size_t)j * nOut, sizeof(double) * nb * KK);
    std::memcpy(h_gram + (size_t)j * KK, pin + (size_t)j * nOut + (wantDots ? nb * KK : 0), sizeof(double) * KK);
  }
  return DH_OK;
}

extern "C" int dh_block_update(dh_context* c, const double* V, int nb, double* W, int64_t dim, int k, const double* coef,
                               const double* m, double* dots, double* gram) {
  if (!c || !V || !W || !coef || !gram || nb <= 0 || dim <= 0) return DH_E_ARG;
  if (k != 1 && k != 2) return dh_fail(c, DH_E_ARG, "the fused update takes 1 or 2 doubles per row");
  int rc = dh_update_segments(c, V, nb, dim * k, W, k, {0}, {dim}, coef, m, dots, gram);
  if (rc == DH_E_RANGE) return dh_fail(c, DH_E_RANGE, "too many basis blocks for the fused update");
  return rc;
}

extern "C" int dh_sectors_block_update(dh_sector_set* S, const double* V, int nb, double* W, int k,
                                       const unsigned char* active, const double* coef, const double* m, double* dots,
                                       double* gram) {
  if (!S || !V || !W || !coef || !gram || nb <= 0) return DH_E_ARG;
  if (k != 1 && k != 2) return dh_fail(S->ctx, DH_E_ARG, "the fused update takes 1 or 2 doubles per row");
  std::vector<int64_t> st, en;
  for (int s = 0; s < S->nbrSectors; ++s)
    if (!active || active[s]) {
      st.push_back(S->offsets[s]);
      en.push_back(S->offsets[s + 1]);
    }
  if (st.empty()) return DH_OK;
  int rc = dh_update_segments(S->ctx, V, nb, S->totalDim * k, W, k, st, en, coef, m, dots, gram);
  if (rc == DH_E_RANGE) return dh_fail(S->ctx, DH_E_RANGE, "too many basis blocks for the fused update");
  return rc;
}
