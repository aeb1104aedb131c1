// Context, device memory and the Lanczos BLAS kernels of libdhgpu.
//
// The Lanczos vector operations replace DiagHam's MultipleRealScalarProductOperation
// (src/Architecture/ArchitectureOperation/MultipleRealScalarProductOperation.cc:174) and
// AddRealLinearCombinationOperation (AddRealLinearCombinationOperation.cc:157) for whole
// Krylov bases at once.  They are HBM-bandwidth kernels: every basis block is read once,
// the working block once (dots) or read+written once (axpy).
#include <cstdio>
#include <cstring>

#include "dh_internal.h"

int dh_fail(dh_context* ctx, int code, const std::string& msg) {
  if (ctx) ctx->lastError = msg;
  return code;
}

extern "C" int dh_create(dh_context** out, int device) {
  if (!out) return DH_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return DH_E_HIP;
  if (device < 0 || device >= n) return DH_E_ARG;
  dh_context* c = new dh_context();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return DH_E_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete c;
    return DH_E_HIP;
  }
  c->nbrCU = prop.multiProcessorCount;
  std::strncpy(c->arch, prop.gcnArchName, sizeof(c->arch) - 1);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return DH_E_HIP;
  }
  c->ownStream = true;
  *out = c;
  return DH_OK;
}

void dh_ctx_retain(dh_context* c) { c->refs.fetch_add(1, std::memory_order_relaxed); }

// the last reference frees the context: by then no object of it is left to use its stream
void dh_ctx_release(dh_context* c) {
  if (c->refs.fetch_sub(1, std::memory_order_acq_rel) != 1) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->d_scratch) hipFree(c->d_scratch);
  if (c->d_pack) hipFree(c->d_pack);
  if (c->h_pinned) hipHostFree(c->h_pinned);
  for (auto& s : c->stage) {
    if (s.host) hipHostFree(s.host);
    if (s.done) hipEventDestroy(s.done);
  }
  if (c->ownStream && c->stream) hipStreamDestroy(c->stream);
  delete c;
}

extern "C" int dh_destroy(dh_context* c) {
  if (!c) return DH_E_ARG;
  if (c->closed) return dh_fail(c, DH_E_STATE, "context already destroyed");
  c->closed = true;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  dh_ctx_release(c);
  return DH_OK;
}

extern "C" const char* dh_last_error(dh_context* c) { return c ? c->lastError.c_str() : "null context"; }

extern "C" int dh_set_stream(dh_context* c, void* s) {
  if (!c) return DH_E_ARG;
  if (c->ownStream && c->stream) {
    hipStreamSynchronize(c->stream);
    hipStreamDestroy(c->stream);
  }
  c->stream = (hipStream_t)s;
  c->ownStream = false;
  return DH_OK;
}

extern "C" int dh_synchronize(dh_context* c) {
  if (!c) return DH_E_ARG;
  DH_HIP(c, hipStreamSynchronize(c->stream));
  return DH_OK;
}

extern "C" int dh_device_info(dh_context* c, int* nbrCU, char* name) {
  if (!c) return DH_E_ARG;
  if (nbrCU) *nbrCU = c->nbrCU;
  if (name) std::strncpy(name, c->arch, 63);
  return DH_OK;
}

extern "C" int dh_memory_info(dh_context* c, size_t* freeBytes, size_t* totalBytes) {
  if (!c || !freeBytes || !totalBytes) return DH_E_ARG;
  hipSetDevice(c->device);
  DH_HIP(c, hipMemGetInfo(freeBytes, totalBytes));
  return DH_OK;
}

static void sample_used(dh_context* c) {
  size_t freeBytes = 0, totalBytes = 0;
  if (hipMemGetInfo(&freeBytes, &totalBytes) == hipSuccess && totalBytes >= freeBytes)
    c->usedPeak = std::max(c->usedPeak, totalBytes - freeBytes);
}

int dh_alloc(dh_context* c, void** p, size_t bytes) {
  hipSetDevice(c->device);
  if (hipMalloc(p, bytes ? bytes : 8) != hipSuccess) return dh_fail(c, DH_E_NOMEM, "hipMalloc failed");
  sample_used(c);
  return DH_OK;
}

extern "C" int dh_memory_peak(dh_context* c, size_t* peakBytes, int reset) {
  if (!c || !peakBytes) return DH_E_ARG;
  hipSetDevice(c->device);
  sample_used(c);
  *peakBytes = c->usedPeak;
  if (reset) {
    c->usedPeak = 0;
    sample_used(c);
  }
  return DH_OK;
}

extern "C" int dh_malloc(dh_context* c, void** p, size_t bytes) {
  if (!c || !p) return DH_E_ARG;
  int rc = dh_alloc(c, p, bytes);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> g(c->allocMutex);
    c->allocs.insert(*p);
  }
  dh_ctx_retain(c);
  return DH_OK;
}

// a buffer keeps its context alive until it is freed, so this may follow dh_destroy
extern "C" int dh_free(dh_context* c, void* p) {
  if (!c) return DH_E_ARG;
  if (!p) return DH_OK;
  {
    std::lock_guard<std::mutex> g(c->allocMutex);
    auto it = c->allocs.find(p);
    if (it == c->allocs.end()) return dh_fail(c, DH_E_ARG, "dh_free: not a live dh_malloc buffer of this context");
    c->allocs.erase(it);
  }
  hipSetDevice(c->device);
  hipError_t e = hipFree(p);
  dh_ctx_release(c);
  if (e != hipSuccess) return DH_E_HIP;
  return DH_OK;
}

extern "C" int dh_memcpy_h2d(dh_context* c, void* dst, const void* src, size_t bytes) {
  if (!c || (!dst && bytes) || (!src && bytes)) return DH_E_ARG;
  DH_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  DH_HIP(c, hipStreamSynchronize(c->stream));
  return DH_OK;
}

extern "C" int dh_memcpy_d2h(dh_context* c, void* dst, const void* src, size_t bytes) {
  if (!c || (!dst && bytes) || (!src && bytes)) return DH_E_ARG;
  DH_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  DH_HIP(c, hipStreamSynchronize(c->stream));
  return DH_OK;
}

__global__ void dh_fill_bytes_kernel(unsigned char* p, unsigned char v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// byte fill by a kernel (no hipMemsetAsync, so callers may capture it in a graph)
extern "C" int dh_memset(dh_context* c, void* dst, int value, size_t bytes) {
  if (!c || (!dst && bytes)) return DH_E_ARG;
  if (!bytes) return DH_OK;
  unsigned blocks = (unsigned)std::min<size_t>((bytes + 255) / 256, 65536);
  hipLaunchKernelGGL(dh_fill_bytes_kernel, dim3(blocks), dim3(256), 0, c->stream, (unsigned char*)dst,
                     (unsigned char)value, bytes);
  DH_HIP(c, hipGetLastError());
  return DH_OK;
}

int dh_scratch(dh_context* c, size_t bytes, double** p) {
  if (bytes > c->scratchBytes) {
    if (c->d_scratch) {
      DH_HIP(c, hipStreamSynchronize(c->stream));
      hipFree(c->d_scratch);
      c->d_scratch = nullptr;
    }
    // grow geometrically: the Krylov basis adds a block per step and would otherwise reallocate often
    size_t nb = std::max<size_t>(std::max<size_t>(bytes, 2 * c->scratchBytes), 1 << 20);
    if (hipMalloc(&c->d_scratch, nb) != hipSuccess) {
      c->scratchBytes = 0;
      return dh_fail(c, DH_E_NOMEM, "scratch allocation failed");
    }
    c->scratchBytes = nb;
  }
  *p = c->d_scratch;
  return DH_OK;
}

int dh_upload_async(dh_context* c, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return DH_OK;
  dh_context::StageSlot& s = c->stage[c->stageNext];
  c->stageNext = (c->stageNext + 1) % dh_context::kStageSlots;
  if (s.pending) {
    DH_HIP(c, hipEventSynchronize(s.done));
    s.pending = false;
  }
  if (bytes > s.cap) {
    if (s.host) hipHostFree(s.host);  // its last copy has completed
    const size_t nb = std::max<size_t>(std::max<size_t>(bytes, 2 * s.cap), 64 << 10);
    s.host = nullptr;
    s.cap = 0;
    if (hipHostMalloc(&s.host, nb, hipHostMallocDefault) != hipSuccess) {
      s.host = nullptr;
      return dh_fail(c, DH_E_NOMEM, "pinned staging allocation failed");
    }
    s.cap = nb;
  }
  if (!s.done) DH_HIP(c, hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
  std::memcpy(s.host, src, bytes);
  DH_HIP(c, hipMemcpyAsync(dst, s.host, bytes, hipMemcpyHostToDevice, c->stream));
  DH_HIP(c, hipEventRecord(s.done, c->stream));
  s.pending = true;
  return DH_OK;
}

int dh_pinned(dh_context* c, size_t bytes, double** p) {
  if (bytes > c->pinnedBytes) {
    if (c->h_pinned) {
      DH_HIP(c, hipStreamSynchronize(c->stream));
      hipHostFree(c->h_pinned);
      c->h_pinned = nullptr;
    }
    size_t nb = std::max<size_t>(std::max<size_t>(bytes, 2 * c->pinnedBytes), 1 << 20);
    if (hipHostMalloc((void**)&c->h_pinned, nb, hipHostMallocDefault) != hipSuccess) {
      c->pinnedBytes = 0;
      return dh_fail(c, DH_E_NOMEM, "pinned allocation failed");
    }
    c->pinnedBytes = nb;
  }
  *p = c->h_pinned;
  return DH_OK;
}

// ----------------------------------------------------------------------------- dots
// partial[(chunk) * total + (b*K + p)*K + q] = sum over this chunk's rows of V_b[row][p] W[row][q]
// G basis blocks per thread pass share one read of W.
template <int K, int G>
__global__ void __launch_bounds__(256) dh_block_dots_kernel(const double* __restrict__ V, int nbrBlocks,
                                                            const double* __restrict__ W, int64_t dim,
                                                            int64_t rowsPerChunk, double* __restrict__ partial,
                                                            bool pairs) {
  constexpr int KK = K * K;
  __shared__ double red[256 / 64][G * KK];  // dh_dots_tile's per-wavefront sums
  const int chunk = blockIdx.y;  // blocks of one chunk are adjacent: its W rows stay in L2
  const int b0 = blockIdx.x * G;
  const int64_t rowBegin = (int64_t)chunk * rowsPerChunk;
  const int64_t rowEnd = min(dim, rowBegin + rowsPerChunk);
  const int64_t total = (int64_t)nbrBlocks * KK;
  dh_dots_tile<K, G>(V, dim * K, nbrBlocks, b0, W, rowBegin, rowEnd, &red[0][0],
                     partial + (int64_t)chunk * total + (int64_t)b0 * KK, pairs);
}

// out[i] = sum over chunks of partial[chunk][i]: a workgroup owns 64 consecutive outputs (one
// per lane, coalesced rows of the partial table); its waves split the chunks and meet in LDS.
__global__ void __launch_bounds__(1024) dh_reduce_chunks_kernel(const double* __restrict__ partial, int nbrChunks,
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

template <int K>
static int launch_dots(dh_context* c, const double* V, int nb, const double* W, int64_t dim, double* h_out) {
  constexpr int G = DotsGroup<K>::value;
  const int KK = K * K;
  // row chunks: enough workgroups in all (about 16 per CU) also when there are few basis blocks to
  // spread over, at least 4 chunks per CU, and no chunk shorter than 1024 rows; an even number of
  // rows, so that the k = 1 row pairs of every chunk are aligned
  const int64_t cols = (nb + G - 1) / G;
  const int64_t want = std::max<int64_t>((int64_t)c->nbrCU * 4, ((int64_t)c->nbrCU * 16 + cols - 1) / cols);
  int64_t target = std::max<int64_t>(1, std::min<int64_t>(want, (dim + 1023) / 1024));
  int64_t rowsPerChunk = ((dim + target - 1) / target + 1) & ~int64_t(1);
  int nbrChunks = (int)((dim + rowsPerChunk - 1) / rowsPerChunk);
  int total = nb * KK;
  double* scratch;
  int rc = dh_scratch(c, sizeof(double) * ((size_t)nbrChunks * total + total), &scratch);
  if (rc) return rc;
  double* d_out = scratch + (size_t)nbrChunks * total;
  dim3 grid((nb + G - 1) / G, nbrChunks);
  hipLaunchKernelGGL((dh_block_dots_kernel<K, G>), grid, dim3(256), 0, c->stream, V, nb, W, dim, rowsPerChunk,
                     scratch, K == 1 && dh_pairs_aligned(V, dim, W));
  hipLaunchKernelGGL(dh_reduce_chunks_kernel, dim3((total + 63) / 64), dim3(1024), 0, c->stream, scratch,
                     nbrChunks, total, d_out);
  DH_HIP(c, hipGetLastError());
  double* pin;
  rc = dh_pinned(c, sizeof(double) * total, &pin);
  if (rc) return rc;
  DH_HIP(c, hipMemcpyAsync(pin, d_out, sizeof(double) * total, hipMemcpyDeviceToHost, c->stream));
  DH_HIP(c, hipStreamSynchronize(c->stream));
  std::memcpy(h_out, pin, sizeof(double) * total);
  return DH_OK;
}

extern "C" int dh_block_dots(dh_context* c, const double* V, int nb, const double* W, int64_t dim, int k,
                             double* out) {
  if (!c || !V || !W || !out || nb <= 0 || dim <= 0) return DH_E_ARG;
  switch (k) {
    case 1: return launch_dots<1>(c, V, nb, W, dim, out);
    case 2: return launch_dots<2>(c, V, nb, W, dim, out);
    case 4: return launch_dots<4>(c, V, nb, W, dim, out);
    case 8: return launch_dots<8>(c, V, nb, W, dim, out);
    case 16: return launch_dots<16>(c, V, nb, W, dim, out);
    default: return dh_fail(c, DH_E_ARG, "k must be 1, 2, 4, 8 or 16");
  }
}

// ----------------------------------------------------------------------------- axpy
// W[row][q] += sum_{b < nb, p} V_b[row][p] C[(b*K+p)*K+q]; C staged in LDS in slabs; at k = 1 the
// basis rows with the nontemporal hint (read once per pass).
template <int K>
__global__ void __launch_bounds__(256) dh_block_axpy_kernel(const double* __restrict__ V, int nb,
                                                            double* __restrict__ W, int64_t dim,
                                                            const double* __restrict__ C) {
  constexpr int KK = K * K;
  constexpr int SLAB = (4096 / KK) > 0 ? (4096 / KK) : 1;  // basis blocks per LDS slab (32 KiB)
  __shared__ double sC[SLAB * KK];
  int64_t row = (int64_t)dh_xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  double acc[K];
#pragma unroll
  for (int q = 0; q < K; ++q) acc[q] = 0.0;
  for (int s0 = 0; s0 < nb; s0 += SLAB) {
    int ns = min(SLAB, nb - s0);
    __syncthreads();
    for (int i = threadIdx.x; i < ns * KK; i += blockDim.x) sC[i] = C[(int64_t)s0 * KK + i];
    __syncthreads();
    if (row < dim) {
      for (int b = 0; b < ns; ++b) {
        const double* vb = V + (int64_t)(s0 + b) * dim * K + row * K;
#pragma unroll
        for (int p = 0; p < K; ++p) {
          const double v = K == 1 ? dh_stream_load(vb + p) : vb[p];
#pragma unroll
          for (int q = 0; q < K; ++q) acc[q] += v * sC[b * KK + p * K + q];
        }
      }
    }
  }
  if (row < dim) {
#pragma unroll
    for (int q = 0; q < K; ++q) W[row * K + q] += acc[q];
  }
}

// k = 1 with at most 8 basis vectors (the local round of a one-vector step): the coefficients go by
// value in the kernel arguments, with no upload; the same products and order as the kernel above
struct SmallCoefficients {
  double c[8];
};
__global__ void __launch_bounds__(256) dh_block_axpy_small_kernel(const double* __restrict__ V, int nb,
                                                                  double* __restrict__ W, int64_t dim,
                                                                  SmallCoefficients C) {
  const int64_t row = (int64_t)dh_xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (row >= dim) return;
  double acc = 0.0;
  for (int b = 0; b < nb; ++b) acc += dh_stream_load(V + (int64_t)b * dim + row) * C.c[b];
  W[row] += acc;
}

template <int K>
static int launch_axpy(dh_context* c, const double* V, int nb, double* W, int64_t dim, const double* h_c) {
  if (K == 1 && nb <= 8) {
    SmallCoefficients C;
    for (int b = 0; b < 8; ++b) C.c[b] = b < nb ? h_c[b] : 0.0;
    hipLaunchKernelGGL(dh_block_axpy_small_kernel, dim3((unsigned)((dim + 255) / 256)), dim3(256), 0, c->stream, V,
                       nb, W, dim, C);
    DH_HIP(c, hipGetLastError());
    return DH_OK;
  }
  const int KK = K * K;
  size_t bytes = sizeof(double) * (size_t)nb * KK;
  double* scratch;
  int rc = dh_scratch(c, bytes, &scratch);
  if (rc) return rc;
  if ((rc = dh_upload_async(c, scratch, h_c, bytes))) return rc;
  unsigned blocks = (unsigned)((dim + 255) / 256);
  hipLaunchKernelGGL((dh_block_axpy_kernel<K>), dim3(blocks), dim3(256), 0, c->stream, V, nb, W, dim, scratch);
  DH_HIP(c, hipGetLastError());
  return DH_OK;
}

extern "C" int dh_block_axpy(dh_context* c, const double* V, int nb, double* W, int64_t dim, int k,
                             const double* coef) {
  if (!c || !V || !W || !coef || nb <= 0 || dim <= 0) return DH_E_ARG;
  switch (k) {
    case 1: return launch_axpy<1>(c, V, nb, W, dim, coef);
    case 2: return launch_axpy<2>(c, V, nb, W, dim, coef);
    case 4: return launch_axpy<4>(c, V, nb, W, dim, coef);
    case 8: return launch_axpy<8>(c, V, nb, W, dim, coef);
    case 16: return launch_axpy<16>(c, V, nb, W, dim, coef);
    default: return dh_fail(c, DH_E_ARG, "k must be 1, 2, 4, 8 or 16");
  }
}

// ----------------------------------------------------------------------------- rmul
template <int K>
__global__ void __launch_bounds__(256) dh_block_rmul_kernel(double* __restrict__ W, int64_t dim,
                                                            const double* __restrict__ M) {
  __shared__ double sM[K * K];
  for (int i = threadIdx.x; i < K * K; i += blockDim.x) sM[i] = M[i];
  __syncthreads();
  int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= dim) return;
  double w[K], o[K];
#pragma unroll
  for (int p = 0; p < K; ++p) w[p] = W[row * K + p];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    double s = 0.0;
#pragma unroll
    for (int p = 0; p < K; ++p) s += w[p] * sM[p * K + q];
    o[q] = s;
  }
#pragma unroll
  for (int q = 0; q < K; ++q) W[row * K + q] = o[q];
}

__global__ void __launch_bounds__(256) dh_block_scale_kernel(double* __restrict__ W, int64_t dim, double m) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= dim) return;
  double s = 0.0;  // the accumulation of dh_block_rmul_kernel<1>, to the bit
  s += W[row] * m;
  W[row] = s;
}

template <int K>
static int launch_rmul(dh_context* c, double* W, int64_t dim, const double* h_m) {
  if (K == 1) {  // a scale: the factor goes by value (W * m, the same product as below)
    hipLaunchKernelGGL(dh_block_scale_kernel, dim3((unsigned)((dim + 255) / 256)), dim3(256), 0, c->stream, W, dim,
                       h_m[0]);
    DH_HIP(c, hipGetLastError());
    return DH_OK;
  }
  size_t bytes = sizeof(double) * K * K;
  double* scratch;
  int rc = dh_scratch(c, bytes, &scratch);
  if (rc) return rc;
  if ((rc = dh_upload_async(c, scratch, h_m, bytes))) return rc;
  hipLaunchKernelGGL((dh_block_rmul_kernel<K>), dim3((unsigned)((dim + 255) / 256)), dim3(256), 0, c->stream, W,
                     dim, scratch);
  DH_HIP(c, hipGetLastError());
  return DH_OK;
}

extern "C" int dh_block_rmul(dh_context* c, double* W, int64_t dim, int k, const double* m) {
  if (!c || !W || !m || dim <= 0) return DH_E_ARG;
  switch (k) {
    case 1: return launch_rmul<1>(c, W, dim, m);
    case 2: return launch_rmul<2>(c, W, dim, m);
    case 4: return launch_rmul<4>(c, W, dim, m);
    case 8: return launch_rmul<8>(c, W, dim, m);
    case 16: return launch_rmul<16>(c, W, dim, m);
    default: return dh_fail(c, DH_E_ARG, "k must be 1, 2, 4, 8 or 16");
  }
}

// ----------------------------------------------------------------------------- axpby / random
__global__ void dh_axpby_kernel(double a, const double* __restrict__ x, double b, double* __restrict__ y,
                                int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] = a * x[i] + b * y[i];
}

extern "C" int dh_axpby(dh_context* c, double a, const double* x, double b, double* y, int64_t n) {
  if (!c || !x || !y || n < 0) return DH_E_ARG;
  if (n == 0) return DH_OK;
  unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, (int64_t)c->nbrCU * 16);
  hipLaunchKernelGGL(dh_axpby_kernel, dim3(blocks), dim3(256), 0, c->stream, a, x, b, y, n);
  DH_HIP(c, hipGetLastError());
  return DH_OK;
}

__device__ __forceinline__ uint64_t dh_splitmix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// value i of the SplitMix64 stream seeded by `seed`, mapped to [-1, 1)
__global__ void dh_fill_random_kernel(double* __restrict__ x, int64_t n, uint64_t seed) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    uint64_t r = dh_splitmix(seed + (uint64_t)(i + 1) * 0x9e3779b97f4a7c15ull);
    x[i] = ((double)(r >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
  }
}

extern "C" int dh_fill_random(dh_context* c, double* x, int64_t n, uint64_t seed) {
  if (!c || !x || n < 0) return DH_E_ARG;
  if (n == 0) return DH_OK;
  unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, (int64_t)c->nbrCU * 16);
  hipLaunchKernelGGL(dh_fill_random_kernel, dim3(blocks), dim3(256), 0, c->stream, x, n, seed);
  DH_HIP(c, hipGetLastError());
  return DH_OK;
}

// ----------------------------------------------------------------------------- block columns
__global__ void dh_get_column_kernel(const double* __restrict__ w, int64_t dim, int k, int col,
                                     double* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < dim; i += stride) out[i] = w[i * k + col];
}

__global__ void dh_set_column_kernel(double* __restrict__ w, int64_t dim, int k, int col,
                                     const double* __restrict__ in) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < dim; i += stride) w[i * k + col] = in[i];
}

extern "C" int dh_block_get_column(dh_context* c, const double* w, int64_t dim, int k, int col, double* out) {
  if (!c || !w || !out || dim < 0 || k <= 0 || col < 0 || col >= k) return DH_E_ARG;
  if (dim == 0) return DH_OK;
  unsigned blocks = (unsigned)std::min<int64_t>((dim + 255) / 256, (int64_t)c->nbrCU * 16);
  hipLaunchKernelGGL(dh_get_column_kernel, dim3(blocks), dim3(256), 0, c->stream, w, dim, k, col, out);
  DH_HIP(c, hipGetLastError());
  return DH_OK;
}

extern "C" int dh_block_set_column(dh_context* c, double* w, int64_t dim, int k, int col, const double* in) {
  if (!c || !w || !in || dim < 0 || k <= 0 || col < 0 || col >= k) return DH_E_ARG;
  if (dim == 0) return DH_OK;
  unsigned blocks = (unsigned)std::min<int64_t>((dim + 255) / 256, (int64_t)c->nbrCU * 16);
  hipLaunchKernelGGL(dh_set_column_kernel, dim3(blocks), dim3(256), 0, c->stream, w, dim, k, col, in);
  DH_HIP(c, hipGetLastError());
  return DH_OK;
}

// ----------------------------------------------------------------------------- copy
// Streaming copy, 16 bytes per lane per iteration (reads n*8 bytes, writes n*8 bytes).
__global__ void __launch_bounds__(256) dh_copy_kernel(double* __restrict__ y, const double* __restrict__ x,
                                                      int64_t n) {
  int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 2;
  int64_t stride = (int64_t)gridDim.x * blockDim.x * 2;
  for (; i + 1 < n; i += stride) {
    double2 v = *reinterpret_cast<const double2*>(x + i);
    *reinterpret_cast<double2*>(y + i) = v;
  }
  if (i < n) y[i] = x[i];
}

extern "C" int dh_copy(dh_context* c, double* y, const double* x, int64_t n) {
  if (!c || !x || !y || n < 0) return DH_E_ARG;
  if (n == 0) return DH_OK;
  if (((uintptr_t)x | (uintptr_t)y) & 15) return dh_fail(c, DH_E_ARG, "dh_copy needs 16-byte aligned vectors");
  unsigned blocks = (unsigned)std::min<int64_t>((n / 2 + 255) / 256 + 1, (int64_t)c->nbrCU * 32);
  hipLaunchKernelGGL(dh_copy_kernel, dim3(blocks), dim3(256), 0, c->stream, y, x, n);
  DH_HIP(c, hipGetLastError());
  return DH_OK;
}
