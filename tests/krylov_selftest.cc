// CPU check of the block Krylov recurrence of the host layer (SegmentedBlockLanczos +
// RunSectorLanczos) on dense matrices held in host memory: the device operations are replaced by
// plain loops over the same block layouts.  Reads a problem description from argv[1] and prints,
// per segment, "converged products e0 e1 ...".
//
// input: complex k maxIter nev precision seed nseg, then nseg dimensions, then per segment the
// dense matrix row-major (real: d*d values; complex: d*d (re, im) pairs)
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <string>
#include <vector>

#include "diagham_host.hpp"

using namespace dhh;

namespace {

struct HostOps : BlockOps {
  std::vector<int64_t> off;
  std::vector<std::vector<double>> H;  // per segment, row-major (complex: interleaved)
  bool cplxMode;
  int k, kr;
  int64_t rows;

  std::vector<int> Active(const unsigned char* active) const {
    std::vector<int> a;
    for (size_t s = 0; s + 1 < off.size(); ++s)
      if (active[s]) a.push_back((int)s);
    return a;
  }
  void Product(const double* X, double* W, const unsigned char* active) override {
    for (int s : Active(active)) {
      const int64_t o = off[s], d = off[s + 1] - off[s];
      for (int64_t r = 0; r < d; ++r)
        for (int q = 0; q < k; ++q) {
          if (!cplxMode) {
            double acc = 0.0;
            for (int64_t c = 0; c < d; ++c) acc += H[s][r * d + c] * X[(o + c) * kr + q];
            W[(o + r) * kr + q] = acc;
          } else {
            double re = 0.0, im = 0.0;
            for (int64_t c = 0; c < d; ++c) {
              const double hr = H[s][2 * (r * d + c)], hi = H[s][2 * (r * d + c) + 1];
              const double xr = X[(o + c) * kr + 2 * q], xi = X[(o + c) * kr + 2 * q + 1];
              re += hr * xr - hi * xi;
              im += hr * xi + hi * xr;
            }
            W[(o + r) * kr + 2 * q] = re;
            W[(o + r) * kr + 2 * q + 1] = im;
          }
        }
    }
  }
  void Dots(const double* V, int nb, const double* W, const unsigned char* active, double* out) override {
    const int KK = kr * kr;
    size_t a = 0;
    for (int s : Active(active)) {
      for (int b = 0; b < nb; ++b)
        for (int p = 0; p < kr; ++p)
          for (int q = 0; q < kr; ++q) {
            double acc = 0.0;
            for (int64_t r = off[s]; r < off[s + 1]; ++r) acc += V[(int64_t)b * rows * kr + r * kr + p] * W[r * kr + q];
            out[a * nb * KK + (size_t)b * KK + p * kr + q] = acc;
          }
      ++a;
    }
  }
  void Axpy(const double* V, int nb, double* W, const unsigned char* active, const double* C) override {
    const int KK = kr * kr;
    size_t a = 0;
    for (int s : Active(active)) {
      for (int64_t r = off[s]; r < off[s + 1]; ++r)
        for (int q = 0; q < kr; ++q) {
          double acc = 0.0;
          for (int b = 0; b < nb; ++b)
            for (int p = 0; p < kr; ++p)
              acc += V[(int64_t)b * rows * kr + r * kr + p] * C[a * nb * KK + (size_t)b * KK + p * kr + q];
          W[r * kr + q] += acc;
        }
      ++a;
    }
  }
  void Rmul(double* W, const unsigned char* active, const double* M) override {
    const int KK = kr * kr;
    size_t a = 0;
    std::vector<double> row(kr);
    for (int s : Active(active)) {
      for (int64_t r = off[s]; r < off[s + 1]; ++r) {
        for (int q = 0; q < kr; ++q) {
          double acc = 0.0;
          for (int p = 0; p < kr; ++p) acc += W[r * kr + p] * M[a * KK + p * kr + q];
          row[q] = acc;
        }
        for (int q = 0; q < kr; ++q) W[r * kr + q] = row[q];
      }
      ++a;
    }
  }
  void Random(double* W, uint64_t seed) override {
    uint64_t x = seed;
    for (int64_t i = 0; i < rows * kr; ++i) {
      uint64_t z = (x += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      W[i] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
  }
};

// grow: the Krylov storage starts with two blocks and grows on demand (SetGrowth), each time by a
// fresh host chunk filled with poison; the output must not change.
template <class T>
void Run(HostOps& ops, int maxIter, int nev, double precision, uint64_t seed, bool grow) {
  std::vector<int64_t> dims;
  for (size_t s = 0; s + 1 < ops.off.size(); ++s) dims.push_back(ops.off[s + 1] - ops.off[s]);
  std::vector<int> caps;
  const int maxBlocks = SectorBlockCaps(dims, ops.k, maxIter, caps);
  const int64_t bd = ops.rows * ops.kr;
  const int first = grow ? std::min(2, maxBlocks) : maxBlocks;
  std::vector<double> storage((size_t)(first + 2) * bd, 0.0);
  SegmentedBlockLanczos<T> lanczos(&ops, dims, ops.k, maxBlocks, storage.data(), bd);
  int growths = 0;
  std::vector<std::vector<double>> chunks;
  if (grow)
    lanczos.SetGrowth(first, [&](int& blocks, int need) {
      if (blocks > need && growths % 2) blocks = need;  // every other growth takes only what is needed
      chunks.emplace_back((size_t)blocks * bd, -7.0);  // poison: every block is written before it is read
      ++growths;
      return chunks.back().data();
    });
  std::vector<SectorResult> out(dims.size());
  std::vector<SectorResult*> res;
  for (auto& r : out) res.push_back(&r);
  RunSectorLanczos(lanczos, caps, nev, precision, seed, res);
  if (grow)
    std::fprintf(stderr, "growths %d capacity %d chunks %d\n", growths, lanczos.BlockCapacity(),
                 lanczos.NbrStorageChunks());
  for (auto& r : out) {
    std::printf("%d %d", r.converged ? 1 : 0, r.products);
    for (double e : r.eigenvalues) std::printf(" %.17g", e);
    std::printf("\n");
  }
}

// Checkpoint round trip: one engine runs `steps` steps; a second runs half of them, saves its host
// state, and a third engine with fresh storage (holding only the saved blocks X_0..X_half) loads
// it and runs the rest.  Prints, per segment, the Ritz values of both and whether every one of
// them is bitwise equal.
template <class T>
void RunResume(HostOps& ops, int maxIter, int nev, double precision, uint64_t seed) {
  std::vector<int64_t> dims;
  for (size_t s = 0; s + 1 < ops.off.size(); ++s) dims.push_back(ops.off[s + 1] - ops.off[s]);
  std::vector<int> caps;
  const int maxBlocks = SectorBlockCaps(dims, ops.k, maxIter, caps);
  const int steps = maxBlocks - 1, half = steps / 2;
  const int64_t bd = ops.rows * ops.kr;
  const std::vector<unsigned char> all(dims.size(), 1);
  auto run = [&](SegmentedBlockLanczos<T>& l, int from, int to) {
    for (int j = from; j < to; ++j) {
      l.Step(all);
      for (size_t s = 0; s < dims.size(); ++s) {
        std::vector<double> cur;
        if (l.CheckDue((int)s, nev)) l.TestConvergence((int)s, nev, precision, cur);
      }
    }
  };
  std::vector<double> st1((size_t)(maxBlocks + 2) * bd), st2(st1.size()), st3(st1.size());
  SegmentedBlockLanczos<T> a(&ops, dims, ops.k, maxBlocks, st1.data(), bd);
  a.Start(seed);
  run(a, 0, steps);
  SegmentedBlockLanczos<T> b(&ops, dims, ops.k, maxBlocks, st2.data(), bd);
  b.Start(seed);
  run(b, 0, half);
  std::stringstream state;
  b.SaveState(state);
  std::copy(st2.begin(), st2.begin() + (size_t)(half + 1) * bd, st3.begin());
  SegmentedBlockLanczos<T> c(&ops, dims, ops.k, maxBlocks, st3.data(), bd);
  c.LoadState(state);
  run(c, half, steps);
  bool same = true;
  for (size_t s = 0; s < dims.size(); ++s) {
    std::vector<double> ea, ec;
    a.Eigen((int)s, ea);
    c.Eigen((int)s, ec);
    same = same && ea.size() == ec.size() && std::equal(ea.begin(), ea.end(), ec.begin());
    std::printf("%zu %zu", ea.size(), ec.size());
    for (size_t i = 0; i < std::min<size_t>(ea.size(), 3); ++i) std::printf(" %.17g %.17g", ea[i], ec[i]);
    std::printf("\n");
  }
  std::printf("%s %d %d\n", same ? "same" : "differ", steps, half);
}

// Orthonormality of the stored Krylov basis after maxIter / k steps: prints the largest
// |<x_i, x_j> - delta_ij| over all stored vectors of every segment (live columns only).
template <class T>
void RunOrtho(HostOps& ops, int maxIter, uint64_t seed) {
  std::vector<int64_t> dims;
  for (size_t s = 0; s + 1 < ops.off.size(); ++s) dims.push_back(ops.off[s + 1] - ops.off[s]);
  std::vector<int> caps;
  const int maxBlocks = SectorBlockCaps(dims, ops.k, maxIter, caps);
  const int steps = maxBlocks - 1;
  const int64_t bd = ops.rows * ops.kr;
  std::vector<double> st((size_t)(maxBlocks + 2) * bd);
  SegmentedBlockLanczos<T> l(&ops, dims, ops.k, maxBlocks, st.data(), bd);
  l.Start(seed);
  const std::vector<unsigned char> all(dims.size(), 1);
  for (int j = 0; j < steps; ++j) l.Step(all);
  double worst = 0.0;
  const int cols = (steps + 1) * ops.k;
  for (size_t s = 0; s < dims.size(); ++s)
    for (int a = 0; a < cols; ++a)
      for (int b = a; b < cols; ++b) {
        double re = 0.0, im = 0.0, na = 0.0;
        const double* va = st.data() + (size_t)(a / ops.k) * bd;
        const double* vb = st.data() + (size_t)(b / ops.k) * bd;
        for (int64_t r = ops.off[s]; r < ops.off[s + 1]; ++r) {
          if (!ops.cplxMode) {
            const double x = va[r * ops.kr + a % ops.k], y = vb[r * ops.kr + b % ops.k];
            re += x * y;
            na += x * x;
          } else {
            const double xr = va[r * ops.kr + 2 * (a % ops.k)], xi = va[r * ops.kr + 2 * (a % ops.k) + 1];
            const double yr = vb[r * ops.kr + 2 * (b % ops.k)], yi = vb[r * ops.kr + 2 * (b % ops.k) + 1];
            re += xr * yr + xi * yi;
            im += xr * yi - xi * yr;
            na += xr * xr + xi * xi;
          }
        }
        if (na == 0.0) continue;  // dead column
        worst = std::max(worst, std::hypot(re - (a == b ? 1.0 : 0.0), im));
      }
  std::printf("%.3e\n", worst);
}

// The split step refuses misuse: finishing a step whose product was not launched, or
// orthogonalising a segment the launched product skipped.  Prints one word per check.
template <class T>
void RunMisuse(HostOps& ops, int maxIter, uint64_t seed) {
  std::vector<int64_t> dims;
  for (size_t s = 0; s + 1 < ops.off.size(); ++s) dims.push_back(ops.off[s + 1] - ops.off[s]);
  std::vector<int> caps;
  const int maxBlocks = SectorBlockCaps(dims, ops.k, maxIter, caps);
  const int64_t bd = ops.rows * ops.kr;
  std::vector<double> st((size_t)(maxBlocks + 2) * bd);
  SegmentedBlockLanczos<T> l(&ops, dims, ops.k, maxBlocks, st.data(), bd);
  l.Start(seed);
  std::vector<unsigned char> all(dims.size(), 1), some(dims.size(), 1);
  some[0] = 0;
  auto refused = [&](auto f) {
    try {
      f();
    } catch (const std::logic_error&) {
      return true;
    }
    return false;
  };
  std::printf("%s\n", refused([&] { l.FinishStep(all); }) ? "refused" : "accepted");
  l.LaunchProduct(some);
  std::printf("%s\n", refused([&] { l.FinishStep(all); }) ? "refused" : "accepted");
  l.LaunchProduct(some);
  std::printf("%s\n", refused([&] { l.FinishStep(some); }) ? "refused" : "accepted");
  std::printf("%s\n", refused([&] { l.FinishStep(some); }) ? "refused" : "accepted");
  std::printf("steps %d\n", l.Steps());
}

}  // namespace

// "lowest" mode: n count, then d[0..n) and e[0..n-1); prints TridiagonalLowest's eigenvalues
static int RunLowest(FILE* f) {
  int n, count;
  if (std::fscanf(f, "%d %d", &n, &count) != 2 || n < 1) return 2;
  std::vector<double> d(n), e(n - 1);
  for (auto& v : d)
    if (std::fscanf(f, "%lf", &v) != 1) return 2;
  for (auto& v : e)
    if (std::fscanf(f, "%lf", &v) != 1) return 2;
  std::vector<double> ev;
  TridiagonalLowest(d, e, count, ev);
  for (double v : ev) std::printf("%.17g\n", v);
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  FILE* f = std::fopen(argv[1], "r");
  if (!f) return 2;
  if (argc > 2 && std::string(argv[2]) == "lowest") return RunLowest(f);
  int cplxMode, k, maxIter, nev, nseg;
  double precision;
  unsigned long long seed;
  if (std::fscanf(f, "%d %d %d %d %lf %llu %d", &cplxMode, &k, &maxIter, &nev, &precision, &seed, &nseg) != 7) return 2;
  HostOps ops;
  ops.cplxMode = cplxMode != 0;
  ops.k = k;
  ops.kr = cplxMode ? 2 * k : k;
  ops.off.assign(1, 0);
  for (int s = 0; s < nseg; ++s) {
    long long d;
    if (std::fscanf(f, "%lld", &d) != 1) return 2;
    ops.off.push_back(ops.off.back() + d);
  }
  ops.rows = ops.off.back();
  ops.H.resize(nseg);
  for (int s = 0; s < nseg; ++s) {
    const int64_t d = ops.off[s + 1] - ops.off[s];
    ops.H[s].resize((size_t)(d * d * (cplxMode ? 2 : 1)));
    for (auto& v : ops.H[s])
      if (std::fscanf(f, "%lf", &v) != 1) return 2;
  }
  std::fclose(f);
  const bool resume = argc > 2 && std::string(argv[2]) == "resume";
  const bool ortho = argc > 2 && std::string(argv[2]) == "ortho";
  const bool grow = argc > 2 && std::string(argv[2]) == "grow";
  const bool misuse = argc > 2 && std::string(argv[2]) == "misuse";
  try {
    if (misuse) {
      if (cplxMode)
        RunMisuse<cplx>(ops, maxIter, seed);
      else
        RunMisuse<double>(ops, maxIter, seed);
      return 0;
    }
    if (ortho) {
      if (cplxMode)
        RunOrtho<cplx>(ops, maxIter, seed);
      else
        RunOrtho<double>(ops, maxIter, seed);
      return 0;
    }
    if (cplxMode)
      resume ? RunResume<cplx>(ops, maxIter, nev, precision, seed) : Run<cplx>(ops, maxIter, nev, precision, seed, grow);
    else
      resume ? RunResume<double>(ops, maxIter, nev, precision, seed)
             : Run<double>(ops, maxIter, nev, precision, seed, grow);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
