// Full-reorthogonalisation (block) Lanczos with a device-resident Krylov basis.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <stdexcept>
#include <string>

#include "diagham_host.hpp"

namespace dhh {

// ------------------------------------------------------------------ small dense eigen tools
// sqrt(a^2 + b^2): the direct form where neither square can overflow or underflow (every
// projected matrix here), std::hypot otherwise; hypot's full scaling made it the main cost of tql2.
static inline double Pythag(double a, double b) {
  const double big = std::max(std::fabs(a), std::fabs(b));
  if (big > 1e150 || (big > 0.0 && big < 1e-150)) return std::hypot(a, b);
  return std::sqrt(a * a + b * b);
}

// Number of eigenvalues of the symmetric tridiagonal (d, e) below x: negative pivots of the
// LDL^T factorisation of T - x (Sturm count); a pivot of magnitude below pivmin is replaced by
// -pivmin, as in LAPACK's dstebz.
static int SturmCount(const std::vector<double>& d, const std::vector<double>& e2, double x, double pivmin) {
  int count = 0;
  double q = 1.0;
  for (size_t i = 0; i < d.size(); ++i) {
    q = d[i] - x - (i > 0 ? e2[i - 1] / q : 0.0);
    if (std::fabs(q) < pivmin) q = -pivmin;
    if (q < 0.0) ++count;
  }
  return count;
}

void TridiagonalLowest(const std::vector<double>& d, const std::vector<double>& e, int count,
                       std::vector<double>& evals, const std::vector<double>* hint) {
  const int n = (int)d.size();
  count = std::max(0, std::min(count, n));
  evals.assign(count, 0.0);
  if (count == 0) return;
  std::vector<double> e2(n > 0 ? n - 1 : 0);
  double lo = d[0], hi = d[0], emax = 0.0;
  for (int i = 0; i < n; ++i) {
    const double left = i > 0 ? std::fabs(e[i - 1]) : 0.0, right = i + 1 < n ? std::fabs(e[i]) : 0.0;
    lo = std::min(lo, d[i] - left - right);
    hi = std::max(hi, d[i] + left + right);
    if (i + 1 < n) {
      e2[i] = e[i] * e[i];
      emax = std::max(emax, e2[i]);
    }
  }
  const double norm = std::max(std::fabs(lo), std::fabs(hi));
  const double pivmin = std::numeric_limits<double>::min() * std::max(1.0, emax);
  lo -= 2.0 * std::numeric_limits<double>::epsilon() * norm + pivmin;
  hi += 2.0 * std::numeric_limits<double>::epsilon() * norm + pivmin;
  // Bisection per eigenvalue down to adjacent doubles: the result is the pair of adjacent doubles
  // where the Sturm count passes i, whatever bracket the bisection starts from.
  std::vector<double> a(count, lo), b(count, hi);
  if (hint)
    for (int i = 0; i < count && i < (int)hint->size(); ++i) {
      // the same eigenvalue of a leading principal submatrix (the tridiagonal matrix of an
      // earlier step) bounds this one from above (Cauchy interlacing), and a converged value moves
      // by rounding only: start from [h - w, h], widening w until the Sturm counts confirm the
      // bracket (62 bisection steps down to adjacent doubles become about 12 for a converged value)
      const double h = (*hint)[i];
      if (!(h > lo && h < hi) || SturmCount(d, e2, h, pivmin) <= i) continue;
      b[i] = h;
      double w = 1e-13 * std::max(1.0, std::fabs(h));
      for (int t = 0; t < 8; ++t, w *= 64.0) {
        const double x = h - w;
        if (x <= lo) break;
        if (SturmCount(d, e2, x, pivmin) <= i) {
          a[i] = x;
          break;
        }
      }
    }
  // the bisections of up to 16 eigenvalues advance together: their Sturm sequences are
  // independent chains of divisions, which the processor overlaps (one chain alone waits on the
  // latency of every division)
  constexpr int kBatch = 16;
  for (int first = 0; first < count; first += kBatch) {
    const int nb = std::min(kBatch, count - first);
    int idx[kBatch];
    double x[kBatch], q[kBatch];
    int cnt[kBatch];
    while (true) {
      int na = 0;
      for (int j = 0; j < nb; ++j) {
        const int i = first + j;
        const double mid = 0.5 * (a[i] + b[i]);
        if (mid <= a[i] || mid >= b[i]) continue;
        idx[na] = i;
        x[na++] = mid;
      }
      if (na == 0) break;
      for (int k = 0; k < na; ++k) {
        q[k] = d[0] - x[k];
        if (std::fabs(q[k]) < pivmin) q[k] = -pivmin;
        cnt[k] = q[k] < 0.0;
      }
      for (int r = 1; r < n; ++r) {
        const double dr = d[r], er = e2[r - 1];
        for (int k = 0; k < na; ++k) {
          double v = dr - x[k] - er / q[k];
          if (std::fabs(v) < pivmin) v = -pivmin;
          q[k] = v;
          cnt[k] += v < 0.0;
        }
      }
      for (int k = 0; k < na; ++k) {
        const int i = idx[k];
        if (cnt[k] > i)
          b[i] = x[k];
        else
          a[i] = x[k];
      }
    }
    for (int j = 0; j < nb; ++j) evals[first + j] = b[first + j];
  }
}

// Implicit-shift QL on a symmetric tridiagonal matrix (the textbook tql2 iteration).
// d: diagonal (n), e: sub-diagonal (e[i] couples i and i+1, n-1 used).  evecs, if given,
// receives the n x n eigenvector matrix column-major (column j <-> evals[j]).
void TridiagonalEigen(int n, std::vector<double> d, std::vector<double> e, std::vector<double>& evals,
                      std::vector<double>* evecs) {
  e.resize(n, 0.0);
  if (n > 0) e[n - 1] = 0.0;
  std::vector<double> z;
  if (evecs) {
    z.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) z[(size_t)i * n + i] = 1.0;  // z[row * n + col]
  }
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    while (true) {
      int m = l;
      for (; m < n - 1; ++m) {
        double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 1e-300 + 2.220446049250313e-16 * dd) break;
      }
      if (m == l) break;
      if (++iter > 60) throw std::runtime_error("TridiagonalEigen: no convergence");
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = Pythag(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
      double s = 1.0, c = 1.0, p = 0.0;
      int i = m - 1;
      bool underflow = false;
      for (; i >= l; --i) {
        double f = s * e[i], b = c * e[i];
        r = Pythag(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          underflow = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        if (evecs) {
          for (int k = 0; k < n; ++k) {
            double zk1 = z[(size_t)k * n + i + 1], zk = z[(size_t)k * n + i];
            z[(size_t)k * n + i + 1] = s * zk + c * zk1;
            z[(size_t)k * n + i] = c * zk - s * zk1;
          }
        }
      }
      if (underflow && i >= l) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return d[a] < d[b]; });
  evals.resize(n);
  for (int i = 0; i < n; ++i) evals[i] = d[order[i]];
  if (evecs) {
    evecs->assign((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < n; ++k) (*evecs)[(size_t)j * n + k] = z[(size_t)k * n + order[j]];
  }
}

// Givens rotation G on the plane (p, p+1) of a symmetric dense matrix: A <- G A G^T,
// restricted to the index window [lo, hi] that holds every nonzero of rows p and p+1.
static inline void rotate_sym(std::vector<double>& a, int n, int p, double c, double s, int lo, int hi) {
  const int q = p + 1;
  double* rp = &a[(size_t)p * n];
  double* rq = &a[(size_t)q * n];
  for (int j = lo; j <= hi; ++j) {
    double x = rp[j], y = rq[j];
    rp[j] = c * x + s * y;
    rq[j] = -s * x + c * y;
  }
  for (int i = lo; i <= hi; ++i) {
    double* ri = &a[(size_t)i * n];
    double x = ri[p], y = ri[q];
    ri[p] = c * x + s * y;
    ri[q] = -s * x + c * y;
  }
}

void BandToTridiagonal(int n, int b, std::vector<double>& a, std::vector<double>& d, std::vector<double>& e,
                       std::vector<double>* q) {
  if (q) {
    q->assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) (*q)[(size_t)i * n + i] = 1.0;  // q[row * n + col]
  }
  auto zero = [&](int p, int col) {
    // annihilate a[p+1][col] against a[p][col] with a rotation in the plane (p, p+1)
    double x = a[(size_t)p * n + col], y = a[(size_t)(p + 1) * n + col];
    if (y == 0.0) return;
    double r = Pythag(x, y);
    double c = x / r, s = y / r;
    int lo = std::max(0, p - b - 1), hi = std::min(n - 1, p + 1 + b + 1);
    rotate_sym(a, n, p, c, s, lo, hi);
    a[(size_t)(p + 1) * n + col] = 0.0;
    a[(size_t)col * n + p + 1] = 0.0;
    if (q) {  // Q <- Q G^T : columns p, p+1
      for (int i = 0; i < n; ++i) {
        double* qi = &(*q)[(size_t)i * n];
        double u = qi[p], v = qi[p + 1];
        qi[p] = c * u + s * v;
        qi[p + 1] = -s * u + c * v;
      }
    }
  };
  if (b > 1) {
    for (int j = 0; j < n - 2; ++j)
      for (int r = std::min(j + b, n - 1); r >= j + 2; --r) {
        zero(r - 1, j);
        // chase the bulge created at (r + b, r - 1)
        for (int p = r; p + b < n; p += b) zero(p + b - 1, p - 1);
      }
  }
  d.resize(n);
  e.assign(n, 0.0);
  for (int i = 0; i < n; ++i) d[i] = a[(size_t)i * n + i];
  for (int i = 0; i + 1 < n; ++i) e[i] = a[(size_t)(i + 1) * n + i];
}

// ------------------------------------------------------------------ block Lanczos
// the block operations of one Hilbert space: its Hamiltonian's product and the block BLAS
struct FullReorthogonalizedBlockLanczosAlgorithm::HamiltonianOps : BlockOps {
  GPUArchitecture* Arch;
  AbstractHamiltonian* Hamiltonian;
  int64_t Dim;
  int K;
  HamiltonianOps(GPUArchitecture* a, AbstractHamiltonian* h, int64_t dim, int k)
      : Arch(a), Hamiltonian(h), Dim(dim), K(k) {}
  void Product(const double* X, double* W, const unsigned char*) override {
    Arch->Check(dh_hv(Hamiltonian->Handle(), X, W, K, 0, 0, -1), "H.v");
  }
  void Dots(const double* V, int nb, const double* W, const unsigned char*, double* out) override {
    Arch->Check(dh_block_dots(Arch->Context(), V, nb, W, Dim, K, out), "block dots");
  }
  void Axpy(const double* V, int nb, double* W, const unsigned char*, const double* C) override {
    Arch->Check(dh_block_axpy(Arch->Context(), V, nb, W, Dim, K, C), "block axpy");
  }
  void Rmul(double* W, const unsigned char*, const double* M) override {
    Arch->Check(dh_block_rmul(Arch->Context(), W, Dim, K, M), "block rmul");
  }
  void Random(double* W, uint64_t seed) override {
    Arch->Check(dh_fill_random(Arch->Context(), W, Dim * K, seed), "random block");
  }
};

FullReorthogonalizedBlockLanczosAlgorithm::FullReorthogonalizedBlockLanczosAlgorithm(
    GPUArchitecture* architecture, AbstractHamiltonian* hamiltonian, int nbrEigenvalue, int blockSize, int maxIter,
    double precision, double residual)
    : Arch(architecture),
      Hamiltonian(hamiltonian),
      NbrEigenvalue(nbrEigenvalue),
      BlockSize(blockSize),
      MaxIter(maxIter),
      Precision(precision),
      Residual(residual) {
  Dim = hamiltonian->GetHilbertSpaceDimension();
  if (BlockSize != 1 && BlockSize != 2 && BlockSize != 4 && BlockSize != 8 && BlockSize != 16)
    throw std::invalid_argument("block size must be 1, 2, 4, 8 or 16");
  // the Krylov space of the sector is complete after at most ceil(Dim / k) blocks
  MaxBlocks = (int)std::min<int64_t>(std::max(3, MaxIter / BlockSize + 1), (Dim + BlockSize - 1) / BlockSize + 1);
  if (!Dense()) {
    const int capacity = InitialKrylovBlocks(BlockSize, MaxBlocks);
    d_basis.push_back(Arch->AllocateMemory((int64_t)(capacity + 2) * Dim * BlockSize));
    Ops.reset(new HamiltonianOps(Arch, Hamiltonian, Dim, BlockSize));
    Engine.reset(new SegmentedBlockLanczos<double>(Ops.get(), std::vector<int64_t>{Dim}, BlockSize, MaxBlocks,
                                                   d_basis[0], Dim * BlockSize));
    Engine->SetGrowth(capacity, [this](int& blocks, int need) {
      return GrowKrylovStorage(Arch, d_basis, blocks, need, Dim * BlockSize);
    });
  }
}

double* GrowKrylovStorage(GPUArchitecture* arch, std::vector<double*>& chunks, int& blocks, int need,
                          int64_t blockDoubles) {
  dh_context* c = arch->Context();
  size_t freeBytes = 0, totalBytes = 0;
  if (dh_memory_info(c, &freeBytes, &totalBytes) == DH_OK) {
    const size_t spare = (size_t)1 << 30, blockBytes = sizeof(double) * (size_t)blockDoubles;
    const int64_t fit = freeBytes > spare ? (int64_t)((freeBytes - spare) / blockBytes) : 0;
    if (fit < need) throw KrylovStorageFull("the device memory cannot hold " + std::to_string(need) + " more Krylov blocks");
    if (blocks > fit) blocks = (int)fit;
  }
  double* p = arch->AllocateMemory((int64_t)blocks * blockDoubles);
  chunks.push_back(p);
  return p;
}

FullReorthogonalizedBlockLanczosAlgorithm::~FullReorthogonalizedBlockLanczosAlgorithm() {
  if (!d_basis.empty()) dh_synchronize(Arch->Context());
  for (double* p : d_basis) Arch->FreeMemory(p);
}

void FullReorthogonalizedBlockLanczosAlgorithm::InitializeLanczosAlgorithm(uint64_t seed) {
  Current.clear();
  if (Dense()) return;
  Engine->Start(seed);
  if (!DiskDirectory.empty()) {
    WriteBlock(0);
    WriteState();
  }
}

void FullReorthogonalizedBlockLanczosAlgorithm::RunLanczosAlgorithm(int nbrIter) {
  const std::vector<unsigned char> all(1, 1);
  for (int it = 0; it < nbrIter; ++it) {
    if (Engine->Steps() > 0 && Engine->Exhausted(0)) return;
    if (Engine->Steps() + 1 > MaxBlocks) return;
    Engine->Step(all);
    if (!DiskDirectory.empty()) {
      WriteBlock(Engine->Steps());
      WriteState();
    }
  }
}

// ------------------------------------------------------------------ disk storage
// lanczos.dat: 8-byte tag, int64 dimension, int32 wanted eigenvalues, double precision, then the
// engine's state (SaveState).  The reference's lanczos.dat (WriteState, :472) holds the
// tridiagonal matrix of its three-vector recurrence; this one holds the block recurrence.
static const char kStateTag[8] = {'D', 'H', 'G', 'P', 'U', 'L', 'Z', '2'};  // 2: with the test cadence

static std::string VectorFileName(const std::string& dir, int64_t index) {
  return dir + "/vector." + std::to_string(index);
}

void FullReorthogonalizedBlockLanczosAlgorithm::SetDiskStorage(const std::string& directory) {
  if (directory.empty()) throw std::invalid_argument("empty directory for the disk storage");
  DiskDirectory = directory;
}

void FullReorthogonalizedBlockLanczosAlgorithm::WriteBlock(int j) {
  const int k = BlockSize;
  double* X = Engine->Block(j);
  if (k == 1) {
    WriteDeviceVector(Arch, VectorFileName(DiskDirectory, j).c_str(), X, Dim, false);
    return;
  }
  double* column = Engine->Scratch();  // the scratch block is free between steps
  for (int p = 0; p < k; ++p) {
    Arch->Check(dh_block_get_column(Arch->Context(), X, Dim, k, p, column), "Krylov column");
    WriteDeviceVector(Arch, VectorFileName(DiskDirectory, (int64_t)j * k + p).c_str(), column, Dim, false);
  }
}

void FullReorthogonalizedBlockLanczosAlgorithm::WriteState() {
  const std::string name = DiskDirectory + "/lanczos.dat";
  const std::string tmp = name + ".tmp";
  {
    std::ofstream o(tmp, std::ios::binary | std::ios::trunc);
    if (!o) throw std::invalid_argument("cannot open the file: " + tmp);
    const int32_t nev = NbrEigenvalue;
    o.write(kStateTag, 8);
    o.write(reinterpret_cast<const char*>(&Dim), 8);
    o.write(reinterpret_cast<const char*>(&nev), 4);
    o.write(reinterpret_cast<const char*>(&Precision), 8);
    Engine->SaveState(o);
    o.flush();
    if (!o) throw std::runtime_error("cannot write " + tmp);
  }
  // a run stopped while writing keeps the previous complete state
  if (std::rename(tmp.c_str(), name.c_str()) != 0) throw std::runtime_error("cannot rename " + tmp);
}

void FullReorthogonalizedBlockLanczosAlgorithm::ResumeLanczosAlgorithm() {
  if (DiskDirectory.empty()) throw std::invalid_argument("resuming needs the directory of the disk storage");
  if (Dense()) throw std::invalid_argument("the space is diagonalised densely: there is no Krylov state to resume");
  const std::string name = DiskDirectory + "/lanczos.dat";
  std::ifstream in(name, std::ios::binary);
  if (!in) throw std::invalid_argument("cannot open the file: " + name);
  char tag[8];
  int64_t dim;
  int32_t nev;
  double precision;
  in.read(tag, 8);
  in.read(reinterpret_cast<char*>(&dim), 8);
  in.read(reinterpret_cast<char*>(&nev), 4);
  in.read(reinterpret_cast<char*>(&precision), 8);
  if (!in || std::memcmp(tag, kStateTag, 8) != 0) throw std::invalid_argument("not a Lanczos state file: " + name);
  if (dim != Dim) throw std::invalid_argument("Hamiltonian does not match dimension of stored vectors in resuming");
  Engine->LoadState(in);
  const int k = BlockSize;
  double* column = Engine->Scratch();
  for (int j = 0; j <= Engine->Steps(); ++j) {
    double* X = Engine->Block(j);
    if (k == 1) {
      ReadDeviceVector(Arch, VectorFileName(DiskDirectory, j).c_str(), X, Dim, false);
      continue;
    }
    for (int p = 0; p < k; ++p) {
      ReadDeviceVector(Arch, VectorFileName(DiskDirectory, (int64_t)j * k + p).c_str(), column, Dim, false);
      Arch->Check(dh_block_set_column(Arch->Context(), X, Dim, k, p, column), "Krylov column");
    }
  }
  Arch->Check(dh_synchronize(Arch->Context()), "sync");
  Current.clear();
}

int FullReorthogonalizedBlockLanczosAlgorithm::NbrBlocks() const { return Engine ? Engine->NbrBlocks(0) : 0; }

void FullReorthogonalizedBlockLanczosAlgorithm::Diagonalize(std::vector<double>& evals, std::vector<double>* evecs) {
  Engine->Eigen(0, evals, evecs);
}

void BlockTridiagonalEigen(const std::vector<std::vector<double>>& A, const std::vector<std::vector<double>>& B,
                           int k, int m, std::vector<double>& evals, std::vector<double>* evecs) {
  const int n = m * k;
  std::vector<double> d, e;
  std::vector<double> Q;
  if (k == 1) {
    d.resize(n);
    e.assign(n, 0.0);
    for (int i = 0; i < m; ++i) {
      d[i] = A[i][0];
      if (i + 1 < m) e[i] = B[i][0];
    }
  } else {
    std::vector<double> T((size_t)n * n, 0.0);
    for (int j = 0; j < m; ++j)
      for (int p = 0; p < k; ++p)
        for (int q = 0; q < k; ++q) {
          T[(size_t)(j * k + p) * n + j * k + q] = A[j][(size_t)p * k + q];
          if (j + 1 < m) {
            double b = B[j][(size_t)p * k + q];  // T[j+1][j] block = B_j
            T[(size_t)((j + 1) * k + p) * n + j * k + q] = b;
            T[(size_t)(j * k + q) * n + (j + 1) * k + p] = b;
          }
        }
    BandToTridiagonal(n, k, T, d, e, evecs ? &Q : nullptr);
  }
  std::vector<double> z;
  TridiagonalEigen(n, d, e, evals, evecs ? &z : nullptr);
  if (evecs) {
    if (k == 1) {
      *evecs = z;
    } else {  // Ritz coefficients in the Krylov basis: Q * z
      evecs->assign((size_t)n * n, 0.0);
      for (int col = 0; col < n; ++col)
        for (int row = 0; row < n; ++row) {
          double s = 0.0;
          for (int t = 0; t < n; ++t) s += Q[(size_t)row * n + t] * z[(size_t)col * n + t];
          (*evecs)[(size_t)col * n + row] = s;
        }
    }
  }
}

// Eigenvector of the symmetric tridiagonal matrix (d, e) for its eigenvalue theta, by inverse
// iteration: two solves of (T - theta I) x = b, each normalised, with T - theta I factored once by
// Gaussian elimination with partial pivoting (row exchanges between neighbouring rows only, so the
// upper factor has two superdiagonals); a pivot that vanishes is replaced by eps |T|.  Costs O(n).
std::vector<double> TridiagonalEigenvector(const std::vector<double>& d, const std::vector<double>& e, double theta) {
  const int n = (int)d.size();
  std::vector<double> diag(n), lower(std::max(n - 1, 0)), up1(std::max(n - 1, 0)), up2(std::max(n - 2, 0));
  std::vector<char> swapped(std::max(n - 1, 0), 0);
  double norm = 0.0;
  for (int i = 0; i < n; ++i) {
    diag[i] = d[i] - theta;
    if (i + 1 < n) lower[i] = up1[i] = e[i];
    norm = std::max(norm, std::fabs(d[i]) + (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(e[i]) : 0.0));
  }
  const double tiny = std::max(norm, 1e-300) * 2.220446049250313e-16;
  for (int i = 0; i + 1 < n; ++i) {
    if (std::fabs(diag[i]) >= std::fabs(lower[i])) {
      if (diag[i] == 0.0) diag[i] = tiny;
      const double f = lower[i] / diag[i];
      lower[i] = f;
      diag[i + 1] -= f * up1[i];
    } else {  // exchange rows i and i + 1
      const double f = diag[i] / lower[i];
      diag[i] = lower[i];
      lower[i] = f;
      const double u = up1[i];
      up1[i] = diag[i + 1];
      diag[i + 1] = u - f * diag[i + 1];
      if (i + 2 < n) {
        up2[i] = up1[i + 1];
        up1[i + 1] = -f * up1[i + 1];
      }
      swapped[i] = 1;
    }
  }
  if (n > 0 && diag[n - 1] == 0.0) diag[n - 1] = tiny;
  std::vector<double> x(n, 1.0);
  for (int it = 0; it < 2; ++it) {
    for (int i = 0; i + 1 < n; ++i) {  // L
      if (swapped[i]) std::swap(x[i], x[i + 1]);
      x[i + 1] -= lower[i] * x[i];
    }
    for (int i = n - 1; i >= 0; --i) {  // U
      double v = x[i];
      if (i + 1 < n) v -= up1[i] * x[i + 1];
      if (i + 2 < n) v -= up2[i] * x[i + 2];
      x[i] = v / diag[i];
    }
    double s = 0.0;
    for (double v : x) s = std::max(s, std::fabs(v));
    for (double& v : x) v /= s;
  }
  double s = 0.0;
  for (double v : x) s += v * v;
  s = std::sqrt(s);
  for (double& v : x) v /= s;
  return x;
}

bool FullReorthogonalizedBlockLanczosAlgorithm::TestConvergence() {
  bool ok = Engine->TestConvergence(0, NbrEigenvalue, Precision, Current);
  const int want = (int)Current.size();
  if (ok && Residual > 0.0 && !Engine->Exhausted(0) && BlockSize == 1) {
    // one vector per step: |H v - theta v| = |beta_m| |last component of the eigenvector of T|,
    // the eigenvectors by inverse iteration (O(m) each; the full eigenvector matrix would cost
    // O(m^3) at every check, minutes for a few thousand steps)
    const int m = Engine->NbrBlocks(0);
    const auto& A = Engine->Diagonal(0);
    const auto& B = Engine->Coupling(0);
    std::vector<double> d(m), e(m > 0 ? m - 1 : 0);
    for (int j = 0; j < m; ++j) {
      d[j] = A[j][0];
      if (j + 1 < m) e[j] = B[j][0];
    }
    std::vector<double> ev;
    Engine->LowestEigen(0, want, ev);
    for (int i = 0; i < want && i < (int)ev.size() && ok; ++i) {
      const std::vector<double> z = TridiagonalEigenvector(d, e, ev[i]);
      if (std::fabs(B[m - 1][0] * z[m - 1]) > Residual * std::fabs(ev[i])) ok = false;
    }
  } else if (ok && Residual > 0.0 && !Engine->Exhausted(0)) {
    // |H v - theta v| = |B_last * (last k coefficients of the Ritz vector)|: the pending
    // block X_m times B[m-1] is what the Krylov space has not yet absorbed
    const int k = BlockSize;
    const int m = Engine->NbrBlocks(0);
    const int n = m * k;
    std::vector<double> ev2, vecs;
    Diagonalize(ev2, &vecs);
    const std::vector<double>& Bl = Engine->Coupling(0)[m - 1];
    for (int i = 0; i < want && ok; ++i) {
      double r2 = 0.0;
      for (int p = 0; p < k; ++p) {
        double s = 0.0;
        for (int q = 0; q < k; ++q) s += Bl[(size_t)p * k + q] * vecs[(size_t)i * n + (m - 1) * k + q];
        r2 += s * s;
      }
      if (std::sqrt(r2) > Residual * std::fabs(ev2[i])) ok = false;
    }
  }
  return ok;
}

void FullReorthogonalizedBlockLanczosAlgorithm::GetEigenstates(int nbr, double* d_out) {
  if (Dense()) throw std::invalid_argument("no Krylov basis: the space was diagonalised densely");
  const int k = BlockSize;
  const int m = Engine->NbrBlocks(0);
  const int n = m * k;
  std::vector<double> ev, vecs;
  Diagonalize(ev, &vecs);
  dh_context* c = Arch->Context();
  nbr = std::min(nbr, (int)ev.size());
  // Ritz vectors in groups of k: block [Dim][k] = V * S, then copied out column by column
  double* tmp = Arch->AllocateMemory(Dim * k);
  for (int g0 = 0; g0 < nbr; g0 += k) {
    Arch->Check(dh_memset(c, tmp, 0, sizeof(double) * Dim * k), "clear");
    std::vector<double> S((size_t)m * k * k, 0.0);
    for (int b = 0; b < m; ++b)
      for (int p = 0; p < k; ++p)
        for (int q = 0; q < k && g0 + q < nbr; ++q) S[((size_t)b * k + p) * k + q] = vecs[(size_t)(g0 + q) * n + b * k + p];
    const unsigned char one = 1;
    Engine->CombineBlocks(0, m, tmp, &one, 1, S.data());
    for (int q = 0; q < k && g0 + q < nbr; ++q)
      Arch->Check(dh_block_get_column(c, tmp, Dim, k, q, d_out + (int64_t)(g0 + q) * Dim), "extract");
  }
  Arch->Check(dh_synchronize(c), "sync");
  Arch->FreeMemory(tmp);
}

// Spaces too small for a Krylov block are diagonalised densely (H built column by column with
// the GPU product), like the reference's full-diagonalisation branch for small dimensions
// (QHEOnSphereMainTask.cc:397).
bool FullReorthogonalizedBlockLanczosAlgorithm::Dense() const { return Dim <= std::max<int64_t>(64, 4 * BlockSize); }

void FullReorthogonalizedBlockLanczosAlgorithm::DenseSolve() {
  dh_context* c = Arch->Context();
  const int n = (int)Dim;
  double* x = Arch->AllocateMemory(n);
  double* y = Arch->AllocateMemory(n);
  std::vector<double> H((size_t)n * n), e(n, 0.0), col(n);
  for (int i = 0; i < n; ++i) {
    e[i] = 1.0;
    Arch->Check(dh_memcpy_h2d(c, x, e.data(), sizeof(double) * n), "unit vector");
    e[i] = 0.0;
    Arch->Check(dh_hv(Hamiltonian->Handle(), x, y, 1, 0, 0, -1), "H.v");
    Arch->Check(dh_memcpy_d2h(c, col.data(), y, sizeof(double) * n), "column");
    for (int r = 0; r < n; ++r) H[(size_t)r * n + i] = col[r];
  }
  Arch->FreeMemory(x);
  Arch->FreeMemory(y);
  for (int r = 0; r < n; ++r)
    for (int q = r + 1; q < n; ++q) H[(size_t)r * n + q] = H[(size_t)q * n + r] = 0.5 * (H[(size_t)r * n + q] + H[(size_t)q * n + r]);
  std::vector<double> d, ee, w;
  BandToTridiagonal(n, std::max(1, n - 1), H, d, ee, nullptr);
  TridiagonalEigen(n, d, ee, w, nullptr);
  Current.assign(w.begin(), w.begin() + std::min<int>(NbrEigenvalue, n));
}

LanczosResult FullReorthogonalizedBlockLanczosAlgorithm::Solve() {
  auto t0 = std::chrono::steady_clock::now();
  LanczosResult res;
  if (Dense()) {
    DenseSolve();
    res.converged = true;
    res.eigenvalues = Current;
    res.iterations = (int)Dim;
    res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return res;
  }
  // a resumed run has already made some of the first steps
  int first = std::max(2, (NbrEigenvalue + 2 + BlockSize - 1) / BlockSize);
  RunLanczosAlgorithm(std::max(0, first - Engine->Steps()));
  const std::vector<unsigned char> all(1, 1);
  while (true) {
    const bool room = !Engine->Exhausted(0) && NbrBlocks() < MaxBlocks;
    // without disk storage the product of the next step runs while the host tests convergence;
    // if the test succeeds that product is discarded
    const bool ahead = room && DiskDirectory.empty();
    if (ahead) Engine->LaunchProduct(all);
    if (Engine->CheckDue(0, NbrEigenvalue) && TestConvergence()) {
      res.converged = true;
      break;
    }
    if (!room) {
      // out of iterations: report the current Ritz values unconverged
      std::vector<double> ev;
      Diagonalize(ev, nullptr);
      Current.assign(ev.begin(), ev.begin() + std::min<int>(NbrEigenvalue, (int)ev.size()));
      break;
    }
    if (ahead)
      Engine->FinishStep(all);
    else
      RunLanczosAlgorithm(1);
  }
  if (!DiskDirectory.empty()) WriteState();  // with the history of the last convergence test
  Arch->Check(dh_synchronize(Arch->Context()), "sync");
  res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  res.eigenvalues = Current;
  res.blockIterations = NbrBlocks();
  res.iterations = NbrBlocks() * BlockSize;
  return res;
}

}  // namespace dhh
