// Fermions on the torus with magnetic translations: complex FP64 H.v over (Kx, Ky) sectors.
//
// Reference path: FermionOnTorusWithMagneticTranslations (FQHE/src/HilbertSpace/
// FermionOnTorusWithMagneticTranslations.cc: GenerateStates :795, AA :473, AdAd :561, and the
// FindCanonicalForm / FindNumberXTranslation / TestXMomentumConstraint helpers of its header) with
// the product of AbstractQHEOnTorusWithMagneticTranslationsHamiltonian.h:160-190.
//
// A sector holds one representative per orbit of the translation T by s = Nphi / gcd(N, Nphi)
// orbitals (the smallest state of the orbit), ascending, restricted to the orbits compatible with
// Kx.  A two-body term maps a representative to a state u; its representative c = T^-t u, the
// coefficient picks up sqrt(orbit(source) / orbit(c)), the fermionic sign of carrying particles
// across the boundary t times (bit t of c's reordering bits) and the phase exp(2 pi i Kx t / g).
// The kernel gathers (H is Hermitian): y[r] += conj(H[t][r]) x[t] over the terms generated from
// row r, like the momentum-space kernel, plus the canonical form and a binary search.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dh_internal.h"

namespace {

struct Translation {
  int shift, comp;
  uint64_t mask;
  __host__ __device__ uint64_t operator()(uint64_t s) const { return (s >> shift) | ((s & mask) << comp); }
};

// the orbit of s: its smallest member, the translations t with T^t(smallest) = s, the orbit size
// and the signature (carried-particle parities summed over one full orbit)
__host__ __device__ inline uint64_t tt_canonical(uint64_t s, const Translation& T, int& nt, int& orbit, int& sig) {
  uint64_t best = s, cur = s;
  int index = 0, at = 0, parity = 0;
  do {
    parity += __builtin_popcountll(cur & T.mask) & 1;
    cur = T(cur);
    ++index;
    if (cur != s && cur < best) {
      best = cur;
      at = index;
    }
  } while (cur != s);
  nt = at ? index - at : 0;
  orbit = index;
  sig = parity;
  return best;
}

__host__ __device__ inline bool tt_kx_allowed(int kx, int orbit, int sig, int modulo, int nbrFermions) {
  if (nbrFermions & 1) return ((kx * orbit) % modulo) == 0;
  return (((kx * orbit) - ((modulo * sig) >> 1)) % modulo) == 0;
}

// representatives of one (Kx, Ky) sector, ascending, with orbit sizes and reordering bits
void build_sector(int N, int nflux, int kx, int ky, const Translation& T, int modulo, std::vector<uint64_t>& states,
                  std::vector<uint8_t>& orbit, std::vector<uint64_t>& reorder) {
  std::vector<uint64_t> raw;
  // N-subsets of total momentum ky mod nflux, highest orbital first (as RawGenerateStates)
  struct Frame {
    int next, left, sum;
    uint64_t st;
  };
  std::vector<Frame> stack;
  stack.push_back({nflux - 1, N, 0, 0ull});
  while (!stack.empty()) {
    Frame f = stack.back();
    stack.pop_back();
    if (f.left == 0) {
      if (f.sum % nflux == ky) {
        int nt, ob, sig;
        raw.push_back(tt_canonical(f.st, T, nt, ob, sig));
      }
      continue;
    }
    if (f.next + 1 < f.left) continue;
    // skip orbital f.next, or occupy it
    stack.push_back({f.next - 1, f.left, f.sum, f.st});
    stack.push_back({f.next - 1, f.left - 1, f.sum + f.next, f.st | (1ull << f.next)});
  }
  std::sort(raw.begin(), raw.end());
  raw.erase(std::unique(raw.begin(), raw.end()), raw.end());
  states.clear();
  orbit.clear();
  reorder.clear();
  for (uint64_t s : raw) {
    int nt, ob, sig;
    tt_canonical(s, T, nt, ob, sig);
    if (!tt_kx_allowed(kx, ob, sig, modulo, N)) continue;
    uint64_t bits = 0;
    if ((N & 1) == 0) {
      uint64_t cur = s;
      for (int k = 1; k <= ob; ++k) {
        const uint64_t moved = (uint64_t)(__builtin_popcountll(cur & T.mask) & 1);
        cur = T(cur);
        bits |= (((bits >> (k - 1)) & 1ull) ^ moved) << k;
      }
    }
    states.push_back(s);
    orbit.push_back((uint8_t)ob);
    reorder.push_back(bits);
  }
}

int gcd(int a, int b) { return b == 0 ? a : gcd(b, a % b); }

}  // namespace

extern "C" int dh_torus_translation_sector_set_create(dh_context* c, dh_sector_set** out, int N, int nflux,
                                                      const int* kx, const int* ky, int nbrSectors) {
  if (!c || !out || !kx || !ky || nbrSectors <= 0) return DH_E_ARG;
  *out = nullptr;
  if (c->closed) return dh_fail(c, DH_E_STATE, "context destroyed");
  if (nflux < 2 || nflux > 62) return dh_fail(c, DH_E_RANGE, "the torus needs 2 to 62 flux quanta");
  if (N <= 0 || N > nflux) return dh_fail(c, DH_E_ARG, "need 0 < nbr_fermions <= nbr_flux");
  hipSetDevice(c->device);
  const int modulo = gcd(N, nflux);
  const Translation T{nflux / modulo, nflux - nflux / modulo, (1ull << (nflux / modulo)) - 1ull};
  dh_sector_set* S = new dh_sector_set();
  S->ctx = c;
  S->kind = DH_SECTORS_TORUS_TRANSLATIONS;
  S->nbrFermions = N;
  S->nbrOrbitals = nflux;
  S->lzMax = nflux;
  S->nbrSiteX = 1;
  S->nbrSiteY = nflux;
  S->nbrSectors = nbrSectors;
  S->modulo = modulo;
  S->stateShift = T.shift;
  S->offsets.assign(nbrSectors + 1, 0);
  std::vector<uint64_t> states, reorder;
  std::vector<uint8_t> orbit;
  std::vector<double> phase;
  for (int i = 0; i < nbrSectors; ++i) {
    if (kx[i] < 0 || kx[i] >= modulo || ky[i] < 0 || ky[i] >= nflux) {
      delete S;
      return dh_fail(c, DH_E_ARG, "(Kx, Ky) outside [0, gcd(N, Nphi)) x [0, Nphi)");
    }
    std::vector<uint64_t> st, ro;
    std::vector<uint8_t> ob;
    build_sector(N, nflux, kx[i], ky[i], T, modulo, st, ob, ro);
    if (st.empty() || st.size() >= (1ull << 31)) {
      delete S;
      return dh_fail(c, DH_E_ARG, "a sector is empty or larger than 2^31");
    }
    states.insert(states.end(), st.begin(), st.end());
    orbit.insert(orbit.end(), ob.begin(), ob.end());
    reorder.insert(reorder.end(), ro.begin(), ro.end());
    S->kx.push_back(kx[i]);
    S->ky.push_back(ky[i]);
    S->offsets[i + 1] = S->offsets[i] + (int64_t)st.size();
    // EvaluateExponentialFactors (AbstractQHEOnTorusWithMagneticTranslationsHamiltonian.cc:67):
    // exp(2 pi i Kx (t s) / Nphi) for t translations of s orbitals
    for (int t = 0; t < modulo; ++t) {
      const double a = 2.0 * M_PI * kx[i] * (double)(t * T.shift) / (double)nflux;
      phase.push_back(std::cos(a));
      phase.push_back(std::sin(a));
    }
  }
  S->totalDim = S->offsets[nbrSectors];
  // top-bits lookup: lut[p] = first row (within the sector) whose state's top lutBits bits are >= p
  int64_t largest = 0;
  for (int i = 0; i < nbrSectors; ++i) largest = std::max(largest, S->offsets[i + 1] - S->offsets[i]);
  int bits = 1;
  while (bits < 16 && bits < nflux && (1ll << bits) < largest) ++bits;
  S->lutBits = bits;
  const int64_t lutSize = (1ll << bits) + 1;
  std::vector<int32_t> lut((size_t)lutSize * nbrSectors);
  for (int i = 0; i < nbrSectors; ++i) {
    const uint64_t* b = states.data() + S->offsets[i];
    const int64_t d = S->offsets[i + 1] - S->offsets[i];
    int64_t row = 0;
    for (int64_t p = 0; p < lutSize; ++p) {
      while (row < d && (b[row] >> (nflux - bits)) < (uint64_t)p) ++row;
      lut[(size_t)i * lutSize + p] = (int32_t)row;
    }
  }
  int rc;
  if ((rc = dh_alloc(c, (void**)&S->d_states, sizeof(uint64_t) * S->totalDim)) ||
      (rc = dh_alloc(c, (void**)&S->d_orbit, S->totalDim)) ||
      (rc = dh_alloc(c, (void**)&S->d_reorder, sizeof(uint64_t) * S->totalDim)) ||
      (rc = dh_alloc(c, (void**)&S->d_lut, sizeof(int32_t) * lut.size())) ||
      (rc = dh_alloc(c, (void**)&S->d_phase, sizeof(double) * phase.size())) ||
      (rc = dh_alloc(c, (void**)&S->d_offsets, sizeof(int64_t) * (nbrSectors + 1)))) {
    dh_sector_set_destroy(S);
    return rc;
  }
  hipMemcpyAsync(S->d_states, states.data(), sizeof(uint64_t) * S->totalDim, hipMemcpyHostToDevice, c->stream);
  hipMemcpyAsync(S->d_orbit, orbit.data(), S->totalDim, hipMemcpyHostToDevice, c->stream);
  hipMemcpyAsync(S->d_reorder, reorder.data(), sizeof(uint64_t) * S->totalDim, hipMemcpyHostToDevice, c->stream);
  hipMemcpyAsync(S->d_lut, lut.data(), sizeof(int32_t) * lut.size(), hipMemcpyHostToDevice, c->stream);
  hipMemcpyAsync(S->d_phase, phase.data(), sizeof(double) * phase.size(), hipMemcpyHostToDevice, c->stream);
  hipMemcpyAsync(S->d_offsets, S->offsets.data(), sizeof(int64_t) * (nbrSectors + 1), hipMemcpyHostToDevice,
                 c->stream);
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    dh_sector_set_destroy(S);
    return dh_fail(c, DH_E_HIP, std::string("torus sector set: ") + hipGetErrorString(e));
  }
  *out = S;
  return DH_OK;
}

extern "C" int dh_sector_set_torus_data(dh_sector_set* S, uint8_t* orbits, uint64_t* reorder) {
  if (!S || S->kind != DH_SECTORS_TORUS_TRANSLATIONS) return DH_E_ARG;
  dh_context* c = S->ctx;
  if (orbits) DH_HIP(c, hipMemcpyAsync(orbits, S->d_orbit, S->totalDim, hipMemcpyDeviceToHost, c->stream));
  if (reorder)
    DH_HIP(c, hipMemcpyAsync(reorder, S->d_reorder, sizeof(uint64_t) * S->totalDim, hipMemcpyDeviceToHost, c->stream));
  DH_HIP(c, hipStreamSynchronize(c->stream));
  return DH_OK;
}

struct TorusArgs {
  const uint64_t* states;
  const uint8_t* orbit;
  const uint64_t* reorder;
  const int32_t* lut;
  const double* phase;
  const int* secKx;  // unused: the phases carry Kx
  const int64_t* offsets;
  int nbrFermions, nflux, modulo, lutBits;
  Translation T;
  int nbrPairs, nbrSums;
  const int* pairInfo;
  const int* sumPairOff;
  const int* sumPairs;
  const int64_t* sumFactorOff;
  const double* factors;
  double shift;
  const int64_t* work;
  int nActive;
  int64_t totalTiles;
};

// index of `st` in the ascending slice b[0..d), or -1
__device__ __forceinline__ int64_t tt_find(const uint64_t* __restrict__ b, const int32_t* __restrict__ lut, int nflux,
                                           int lutBits, uint64_t st) {
  const uint64_t p = st >> (nflux - lutBits);
  int64_t lo = lut[p], hi = lut[p + 1];
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (b[mid] < st)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < lut[p + 1] && b[lo] == st) ? lo : -1;
}

template <int K, int LPR>
__global__ void __launch_bounds__(1024) torus_translation_hv_kernel(TorusArgs a, const double* __restrict__ x,
                                                                    double* __restrict__ y, int add) {
  extern __shared__ int64_t smem64[];
  int64_t* sWork = smem64;                                        // 2 nActive + 1
  int64_t* sFacOff = sWork + 2 * a.nActive + 1;                   // nbrSums + 1
  int* sInfo = reinterpret_cast<int*>(sFacOff + a.nbrSums + 1);  // nbrPairs
  int* sPairOff = sInfo + a.nbrPairs;                             // nbrSums + 1
  int* sPairs = sPairOff + a.nbrSums + 1;                         // nbrPairs
  for (int i = threadIdx.x; i < 2 * a.nActive + 1; i += blockDim.x) sWork[i] = a.work[i];
  for (int i = threadIdx.x; i <= a.nbrSums; i += blockDim.x) {
    sFacOff[i] = a.sumFactorOff[i];
    sPairOff[i] = a.sumPairOff[i];
  }
  for (int i = threadIdx.x; i < a.nbrPairs; i += blockDim.x) {
    sInfo[i] = a.pairInfo[i];
    sPairs[i] = a.sumPairs[i];
  }
  __syncthreads();
  const int64_t* sPrefix = sWork + a.nActive;
  const int lane = threadIdx.x % LPR;
  const int64_t rowsPerTile = blockDim.x / LPR;
  const int64_t lutSize = (1ll << a.lutBits) + 1;
  for (int64_t tile = dh_xcd_block(blockIdx.x, gridDim.x); tile < a.totalTiles; tile += gridDim.x) {
    int lo = 0, hi = a.nActive;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (sPrefix[mid] <= tile)
        lo = mid;
      else
        hi = mid;
    }
    const int sec = (int)sWork[lo];
    const int64_t off = a.offsets[sec];
    const int64_t dim = a.offsets[sec + 1] - off;
    const uint64_t* b = a.states + off;
    const int32_t* lut = a.lut + (int64_t)sec * lutSize;
    const double* ph = a.phase + 2 * (int64_t)sec * a.modulo;
    const int64_t local = (tile - sPrefix[lo]) * rowsPerTile + threadIdx.x / LPR;
    const bool active = local < dim;
    const int64_t g = off + (active ? local : 0);
    const uint64_t s = active ? a.states[g] : 0ull;
    const double orbitRow = active ? (double)a.orbit[g] : 1.0;
    double acc[2 * K];
#pragma unroll
    for (int v = 0; v < 2 * K; ++v) acc[v] = 0.0;
    const double* xs = x + off * 2 * K;
    for (uint64_t r1 = s; r1; r1 &= r1 - 1) {
      const int n1 = __builtin_ctzll(r1);  // lower annihilated orbital
      for (uint64_t r2 = s & ~((2ull << n1) - 1ull); r2; r2 &= r2 - 1) {
        const int n2 = __builtin_ctzll(r2);  // higher annihilated orbital
        // AA(n1, n2): a_n2 first, then a_n1 (sign: occupied orbitals at and above each)
        const int sgnA = __popcll(s >> n2) + __popcll((s ^ (1ull << n2)) >> n1);
        const uint64_t t = s ^ (1ull << n1) ^ (1ull << n2);
        const int info = sInfo[(n2 * (n2 - 1)) / 2 + n1];
        const int sum = info >> 16, i1 = info & 0xffff;
        const int p0 = sPairOff[sum], np = sPairOff[sum + 1] - p0;
        const double* F = a.factors + 2 * (sFacOff[sum] + (int64_t)i1 * np);
        for (int i2 = lane; i2 < np; i2 += LPR) {
          const int pk = sPairs[p0 + i2];
          const int m1 = pk & 0xff, m2 = pk >> 8;  // m1 < m2
          const uint64_t cm = (1ull << m1) | (1ull << m2);
          if (t & cm) continue;
          const double fr = F[2 * i2], fi = F[2 * i2 + 1];
          if (fr == 0.0 && fi == 0.0) continue;
          // AdAd(m1, m2): a+_m2 first, then a+_m1 (sign: occupied orbitals above each)
          const int sg = sgnA + __popcll(t >> m2) + __popcll((t | (1ull << m2)) >> m1);
          int nt, orb, sig;
          const uint64_t canon = tt_canonical(t | cm, a.T, nt, orb, sig);
          // an orbit outside the Kx constraint has no representative in the sector
          const int64_t idx = tt_find(b, lut, a.nflux, a.lutBits, canon);
          if (idx < 0) continue;
          double scale = sqrt(orbitRow / (double)a.orbit[off + idx]);
          if (((a.reorder[off + idx] >> nt) & 1ull) ^ (uint64_t)(sg & 1)) scale = -scale;
          // H[target][row] = scale F E(nt); the gather takes its conjugate
          const double er = ph[2 * nt], ei = ph[2 * nt + 1];
          const double hr = scale * (fr * er - fi * ei), hi = scale * (fr * ei + fi * er);
          const double cr = hr, ci = -hi;
          const double* xr = xs + idx * 2 * K;
#pragma unroll
          for (int v = 0; v < K; ++v) {
            double2 p = *reinterpret_cast<const double2*>(xr + 2 * v);
            acc[2 * v] += cr * p.x - ci * p.y;
            acc[2 * v + 1] += cr * p.y + ci * p.x;
          }
        }
      }
    }
    if constexpr (LPR > 1) {
#pragma unroll
      for (int v = 0; v < 2 * K; ++v)
        for (int o = LPR / 2; o > 0; o >>= 1) acc[v] += __shfl_xor(acc[v], o, LPR);
    }
    if (active && lane == 0) {
      const double* xr = x + g * 2 * K;
      double* yr = y + g * 2 * K;
#pragma unroll
      for (int v = 0; v < 2 * K; ++v) yr[v] = (add ? yr[v] : 0.0) + acc[v] + a.shift * xr[v];
    }
  }
}

template <int K, int LPR>
static int torus_launch(dh_sector_set* S, const double* x, double* y, int add, const unsigned char* active) {
  dh_context* c = S->ctx;
  const int64_t rowsPerTile = S->blockSize / LPR;
  int na;
  int64_t tiles;
  int rc = dh_sectors_upload_work(S, active, rowsPerTile, &na, &tiles);
  if (rc) return rc;
  if (tiles == 0) return DH_OK;
  TorusArgs a;
  a.states = S->d_states;
  a.orbit = S->d_orbit;
  a.reorder = S->d_reorder;
  a.lut = S->d_lut;
  a.phase = S->d_phase;
  a.secKx = nullptr;
  a.offsets = S->d_offsets;
  a.nbrFermions = S->nbrFermions;
  a.nflux = S->nbrOrbitals;
  a.modulo = S->modulo;
  a.lutBits = S->lutBits;
  a.T = Translation{S->stateShift, S->nbrOrbitals - S->stateShift, (1ull << S->stateShift) - 1ull};
  a.nbrPairs = S->nbrPairs;
  a.nbrSums = S->nbrSums;
  a.pairInfo = S->d_pairInfo;
  a.sumPairOff = S->d_sumPairOff;
  a.sumPairs = S->d_sumPairs;
  a.sumFactorOff = S->d_sumFactorOff;
  a.factors = S->d_cfactors;
  a.shift = S->shift;
  a.work = S->d_work;
  a.nActive = na;
  a.totalTiles = tiles;
  const size_t lds = 8 * (2 * (size_t)na + 1) + 8 * (size_t)(S->nbrSums + 1) + 4 * (size_t)S->nbrPairs +
                     4 * (size_t)(S->nbrSums + 1) + 4 * (size_t)S->nbrPairs;
  if (lds > 64 * 1024) return dh_fail(c, DH_E_RANGE, "staged tables exceed LDS");
  auto kern = torus_translation_hv_kernel<K, LPR>;
  const int perCU = dh_blocks_per_cu(c, kern, S->blockSize, lds);
  const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(tiles, (int64_t)c->nbrCU * perCU));
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(S->blockSize), lds, c->stream, a, x, y, add);
  DH_HIP(c, hipGetLastError());
  return DH_OK;
}

template <int K>
static int torus_lpr(dh_sector_set* S, const double* x, double* y, int add, const unsigned char* active) {
  switch (S->lanesPerRow) {
    case 1: return torus_launch<K, 1>(S, x, y, add, active);
    case 2: return torus_launch<K, 2>(S, x, y, add, active);
    case 4: return torus_launch<K, 4>(S, x, y, add, active);
    default: return torus_launch<K, 8>(S, x, y, add, active);
  }
}

int dh_torus_translation_sectors_hv(dh_sector_set* S, const double* x, double* y, int k, int add,
                                    const unsigned char* active) {
  if (!S->d_cfactors) return dh_fail(S->ctx, DH_E_STATE, "torus factors not set");
  switch (k) {
    case 1: return torus_lpr<1>(S, x, y, add, active);
    case 2: return torus_lpr<2>(S, x, y, add, active);
    case 4: return torus_lpr<4>(S, x, y, add, active);
    case 8: return torus_lpr<8>(S, x, y, add, active);
    default: return dh_fail(S->ctx, DH_E_ARG, "k must be 1, 2, 4 or 8 complex vectors");
  }
}
