// Block dot products that allocate nothing: for callers that cannot call hipMalloc (code under HIP
// graph capture, tight serving loops, harnesses that forbid allocation).  The row-chunk partial sums
// live in a file-scope device array of fixed size instead of the context's growable scratch, and the
// result stays in device memory (no pinned staging, no wait).
#include <mutex>

#include "dh_internal.h"

namespace {
// partial sums of the row chunks: at most kScratchDoubles (4 MiB) per device
constexpr int64_t kScratchDoubles = (int64_t)1 << 19;
constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
double* g_scratchPtr[kMaxDevices] = {nullptr};
}  // namespace

// One copy per device, kept until the process ends.  Host code never passes this name to a kernel:
// it passes the device address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double scratch[kScratchDoubles];

template <int K, int G>
__global__ void __launch_bounds__(256) noalloc_dots_kernel(const double* __restrict__ V, int nbrBlocks,
                                                           const double* __restrict__ W, int64_t dim,
                                                           int64_t rowsPerChunk, double* __restrict__ partial,
                                                           bool pairs) {
  constexpr int KK = K * K;
  __shared__ double red[256 / 64][G * KK];
  const int chunk = blockIdx.y;
  const int b0 = blockIdx.x * G;
  const int64_t rowBegin = (int64_t)chunk * rowsPerChunk;
  const int64_t rowEnd = min(dim, rowBegin + rowsPerChunk);
  const int64_t total = (int64_t)nbrBlocks * KK;
  dh_dots_tile<K, G>(V, dim * K, nbrBlocks, b0, W, rowBegin, rowEnd, &red[0][0],
                     partial + (int64_t)chunk * total + (int64_t)b0 * KK, pairs);
}

// out[i] = sum over the chunks of partial[chunk][i] (64 outputs per workgroup, waves split the chunks)
__global__ void __launch_bounds__(1024) noalloc_reduce_kernel(const double* __restrict__ partial, int nbrChunks,
                                                              int total, double* __restrict__ out) {
  __shared__ double red[16][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int i = blockIdx.x * 64 + lane;
  double s = 0.0;
  if (i < total)
    for (int c = wave; c < nbrChunks; c += nw) s += partial[(int64_t)c * total + i];
  red[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && i < total) {
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += red[w][lane];
    out[i] = t;
  }
}

// device address of this device's scratch array, looked up on the first call on the device
static int scratch_address(dh_context* c, double** p) {
  if (c->device < 0 || c->device >= kMaxDevices) return dh_fail(c, DH_E_RANGE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> g(g_scratchMutex);
  if (!g_scratchPtr[c->device]) {
    DH_HIP(c, hipSetDevice(c->device));
    void* ptr = nullptr;
    DH_HIP(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return dh_fail(c, DH_E_HIP, "no device address for the dots scratch array");
    g_scratchPtr[c->device] = (double*)ptr;
  }
  *p = g_scratchPtr[c->device];
  return DH_OK;
}

template <int K>
static int launch_noalloc_dots(dh_context* c, const double* V, int nb, const double* W, int64_t dim, double* d_out) {
  constexpr int G = DotsGroup<K>::value;
  const int64_t total = (int64_t)nb * K * K;
  if (total > kScratchDoubles)
    return dh_fail(c, DH_E_RANGE, "dh_block_dots_noalloc: nbr_blocks * k * k exceeds the 524288 scratch doubles");
  // the row chunks of dh_block_dots, fewer when their partial sums would not fit the scratch array
  const int64_t cols = (nb + G - 1) / G;
  const int64_t want = std::max<int64_t>((int64_t)c->nbrCU * 4, ((int64_t)c->nbrCU * 16 + cols - 1) / cols);
  int64_t target = std::max<int64_t>(1, std::min<int64_t>(want, (dim + 1023) / 1024));
  target = std::min<int64_t>(target, kScratchDoubles / total);
  const int64_t rowsPerChunk = ((dim + target - 1) / target + 1) & ~int64_t(1);
  const int nbrChunks = (int)((dim + rowsPerChunk - 1) / rowsPerChunk);
  double* partial = nullptr;
  int rc = scratch_address(c, &partial);
  if (rc) return rc;
  dim3 grid((unsigned)cols, (unsigned)nbrChunks);
  hipLaunchKernelGGL((noalloc_dots_kernel<K, G>), grid, dim3(256), 0, c->stream, V, nb, W, dim, rowsPerChunk, partial,
                     K == 1 && dh_pairs_aligned(V, dim, W));
  hipLaunchKernelGGL(noalloc_reduce_kernel, dim3((unsigned)((total + 63) / 64)), dim3(1024), 0, c->stream, partial,
                     nbrChunks, (int)total, d_out);
  DH_HIP(c, hipGetLastError());
  return DH_OK;
}

extern "C" int dh_block_dots_noalloc(dh_context* c, const double* V, int nb, const double* W, int64_t dim, int k,
                                     double* d_out) {
  if (!c || !V || !W || !d_out || nb <= 0 || dim <= 0) return DH_E_ARG;
  switch (k) {
    case 1: return launch_noalloc_dots<1>(c, V, nb, W, dim, d_out);
    case 2: return launch_noalloc_dots<2>(c, V, nb, W, dim, d_out);
    case 4: return launch_noalloc_dots<4>(c, V, nb, W, dim, d_out);
    case 8: return launch_noalloc_dots<8>(c, V, nb, W, dim, d_out);
    case 16: return launch_noalloc_dots<16>(c, V, nb, W, dim, d_out);
    default: return dh_fail(c, DH_E_ARG, "k must be 1, 2, 4, 8 or 16");
  }
}
