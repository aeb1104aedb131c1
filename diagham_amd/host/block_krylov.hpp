// Block Lanczos recurrence with full reorthogonalisation, shared by the single-space algorithm
// (FullReorthogonalizedBlockLanczosAlgorithm.cc) and the drivers that advance many sectors
// together.  The Krylov basis stays on the device; the host keeps the small k x k blocks.
//
// Orthogonalisation.  The new block W = H X_j goes through two rounds of "project out the basis,
// then Cholesky-QR of the block" (block classical Gram-Schmidt with reorthogonalisation).  The
// second projection removes what the first Cholesky-QR amplified when the columns of W were
// nearly dependent, keeps the basis orthonormal to rounding level (which the 1e-14 relative
// convergence test needs), and exposes columns that were only noise.
//
// Rank deficiency.  A column is deficient when its Cholesky pivot is negligible against its own
// norm (it depends on the previous columns of the block), or when a projection leaves only
// rounding noise of what it had before (it lies in the span of the basis).  Such a column carries
// no new direction: either the Krylov space of the start block became invariant, or the sector
// has fewer than k directions left.  It gets a random vector orthogonalised against the basis and
// the rest of the block (the restart branch of FullReorthogonalizedLanczosAlgorithm.cc:270).
// When no direction is left it is zeroed and marked dead: it stays decoupled from the block
// tridiagonal matrix (zero rows in the coupling blocks) and its diagonal entry is replaced by a
// value above the Gershgorin bound of the matrix, so it never appears among the wanted
// eigenvalues.  A block whose columns are all dead means the Krylov space of the sector is
// complete and the projected matrix is exact.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <istream>
#include <ostream>
#include <stdexcept>
#include <vector>

namespace dhh {

using cplx = std::complex<double>;

// Device operations of the recurrence.  A block is `rows` x KR doubles, row-major (a block of k
// complex vectors interleaves real and imaginary parts: KR = 2k).  Rows are split in segments
// (sectors); an operation touches only the segments flagged in `active`, and the per-segment host
// matrices (KR x KR each) are packed for the active segments in increasing order.
struct BlockOps {
  virtual ~BlockOps() {}
  virtual void Product(const double* X, double* W, const unsigned char* active) = 0;  // W = H X
  // out[a][b] = V_b^T W over segment a (b < nb)
  virtual void Dots(const double* V, int nb, const double* W, const unsigned char* active, double* out) = 0;
  // W += sum_b V_b C[a][b] over segment a
  virtual void Axpy(const double* V, int nb, double* W, const unsigned char* active, const double* C) = 0;
  // W <- W M[a] over segment a
  virtual void Rmul(double* W, const unsigned char* active, const double* M) = 0;
  virtual void Random(double* W, uint64_t seed) = 0;  // the whole block
  // W <- W M[a] + sum_b V_b C[a][b] (M null: W M = W), then dots[a][b] = V_b^T W (dots may be null)
  // and gram[a] = W^T W over the updated W.  Implementations may do it in one pass over V; this
  // default composes the operations above.
  virtual void Update(const double* V, int nb, double* W, const unsigned char* active, const double* C,
                      const double* M, double* dots, double* gram) {
    if (M) Rmul(W, active, M);
    Axpy(V, nb, W, active, C);
    if (dots) Dots(V, nb, W, active, dots);
    Dots(W, 1, W, active, gram);
  }
};

template <class T>
struct KrylovScalar;

template <>
struct KrylovScalar<double> {
  static constexpr int Real = 1;
  static double Norm(double x) { return x * x; }
  static double Conj(double x) { return x; }
  // k x k coefficient block from the real Gram block of the device layout, and back
  static void FromGram(int k, const double* G, double* C) { std::memcpy(C, G, sizeof(double) * k * k); }
  static void ToDevice(int k, const double* C, double* M) { std::memcpy(M, C, sizeof(double) * k * k); }
};

template <>
struct KrylovScalar<cplx> {
  static constexpr int Real = 2;
  static double Norm(cplx x) { return std::norm(x); }
  static cplx Conj(cplx x) { return std::conj(x); }
  // C[p][q] = <v_p, w_q> from the real (2k x 2k) Gram block of interleaved rows
  static void FromGram(int k, const double* G, cplx* C) {
    const int K2 = 2 * k;
    for (int p = 0; p < k; ++p)
      for (int q = 0; q < k; ++q) {
        double rr = G[(size_t)(2 * p) * K2 + 2 * q], ii = G[(size_t)(2 * p + 1) * K2 + 2 * q + 1];
        double ri = G[(size_t)(2 * p) * K2 + 2 * q + 1], ir = G[(size_t)(2 * p + 1) * K2 + 2 * q];
        C[(size_t)p * k + q] = cplx(rr + ii, ri - ir);
      }
  }
  // real (2k x 2k) matrix acting on interleaved rows: w_q += sum_p v_p c_pq
  static void ToDevice(int k, const cplx* C, double* M) {
    const int K2 = 2 * k;
    for (int p = 0; p < k; ++p)
      for (int q = 0; q < k; ++q) {
        cplx c = C[(size_t)p * k + q];
        M[(size_t)(2 * p) * K2 + 2 * q] = c.real();
        M[(size_t)(2 * p) * K2 + 2 * q + 1] = c.imag();
        M[(size_t)(2 * p + 1) * K2 + 2 * q] = -c.imag();
        M[(size_t)(2 * p + 1) * K2 + 2 * q + 1] = c.real();
      }
  }
};

// Eigenvalues (ascending) of the (Hermitian) block tridiagonal matrix with diagonal blocks
// A[0..m) and sub-diagonal blocks B[0..m-1); Ritz coefficients (column-major n x n) for real
// matrices only.
void BlockTridiagonalEigen(const std::vector<std::vector<double>>& A, const std::vector<std::vector<double>>& B,
                           int k, int m, std::vector<double>& evals, std::vector<double>* evecs);
void BlockTridiagonalEigen(const std::vector<std::vector<cplx>>& A, const std::vector<std::vector<cplx>>& B, int k,
                           int m, std::vector<double>& evals, std::vector<cplx>* evecs);
// The `count` lowest eigenvalues (ascending) of the symmetric tridiagonal matrix with diagonal d
// and off-diagonal e, by Sturm-count bisection to adjacent doubles: O(count * n * 60) instead of
// the O(n^2) of a full QL sweep, for the per-step convergence tests of the one-vector recurrence.
// hint: the lowest eigenvalues of a leading principal submatrix (an earlier step's matrix), which
// narrow the starting brackets; the results are the same with or without it.
void TridiagonalLowest(const std::vector<double>& d, const std::vector<double>& e, int count,
                       std::vector<double>& evals, const std::vector<double>* hint = nullptr);

template <class T>
class SegmentedBlockLanczos {
 public:
  using S = KrylovScalar<T>;

  // storage: (maxBlocks + 2) device blocks of blockDoubles (the last one is scratch), or fewer with
  // SetGrowth
  SegmentedBlockLanczos(BlockOps* ops, const std::vector<int64_t>& dims, int k, int maxBlocks, double* storage,
                        int64_t blockDoubles)
      : Ops(ops),
        Dims(dims),
        K(k),
        KR(k * S::Real),
        MaxBlocks(maxBlocks),
        Capacity(maxBlocks),
        ScratchBlock(storage + (int64_t)(maxBlocks + 1) * blockDoubles),
        BlockDoubles(blockDoubles),
        A(dims.size()),
        B(dims.size()),
        DeadCols(dims.size()),
        Previous(dims.size()),
        LastCheck(dims.size(), -1),
        LastChange(dims.size(), kFarChange),
        LowestHint(dims.size()) {
    Chunks.push_back({storage, 0, maxBlocks + 1});
  }

  // Krylov block j (0 <= j <= BlockCapacity())
  double* Block(int j) const {
    for (const Chunk& c : Chunks)
      if (j < c.first + c.count) return c.ptr + (int64_t)(j - c.first) * BlockDoubles;
    throw std::logic_error("Krylov block outside the storage");
  }
  // the scratch block (free between steps)
  double* Scratch() const { return ScratchBlock; }
  int BlockCapacity() const { return Capacity; }
  int NbrStorageChunks() const { return (int)Chunks.size(); }
  // W += sum_b X_(b0 + b) C[a][b] over segment a of `mask` (na of them), in device layout
  void CombineBlocks(int b0, int nb, double* W, const unsigned char* mask, size_t na, const double* C) {
    BasisAxpy(b0, nb, W, mask, na, C);
  }

  // Krylov storage that grows on demand instead of holding maxBlocks + 2 blocks from the start: the
  // storage passed to the constructor holds `capacity` + 2 blocks (blocks 0 .. capacity, then the
  // scratch block), and grow(blocks, need) must return storage for `blocks` more blocks, which it
  // may lower to what fits, but not below `need`.  The new blocks are a chunk of their own: nothing
  // is copied, so the old and the new storage never need room together.  Each growth adds at
  // least as many blocks as the storage holds.
  using GrowFn = std::function<double*(int& blocks, int need)>;
  void SetGrowth(int capacity, GrowFn grow) {
    Capacity = std::max(1, std::min(capacity, MaxBlocks));
    ScratchBlock = Chunks[0].ptr + (int64_t)(Capacity + 1) * BlockDoubles;
    Chunks[0].count = Capacity + 1;
    Grow = std::move(grow);
  }
  // room for the Krylov blocks X_0 .. X_blocks
  void Reserve(int blocks) {
    if (blocks <= Capacity) return;
    if (blocks > MaxBlocks || !Grow) throw std::runtime_error("Krylov storage exhausted");
    const int need = blocks - Capacity;
    int more = std::min(MaxBlocks - Capacity, std::max(need, Capacity + 1));
    double* p = Grow(more, need);
    if (more < need || Capacity + more > MaxBlocks) throw std::logic_error("Krylov storage grown outside its bounds");
    Chunks.push_back({p, Capacity + 1, more});
    Capacity += more;
  }

  int BlockSize() const { return K; }
  int Steps() const { return StepCount; }
  int NbrBlocks(int s) const { return (int)A[s].size(); }
  const std::vector<std::vector<T>>& Diagonal(int s) const { return A[s]; }
  const std::vector<std::vector<T>>& Coupling(int s) const { return B[s]; }

  // random orthonormal start block X_0 in every segment
  void Start(uint64_t seed) {
    NextSeed = seed * 0x9E3779B97F4A7C15ull + 0x5EED;
    Ops->Random(Block(0), seed);
    std::vector<int> act(Dims.size());
    for (size_t s = 0; s < Dims.size(); ++s) act[s] = (int)s;
    std::vector<Rounds> r;
    Orthonormalize(Block(0), 0, act, r);
    for (size_t s = 0; s < Dims.size(); ++s) {
      A[s].clear();
      B[s].clear();
      DeadCols[s].assign(1, Dead(r[s].good));
      Previous[s].clear();
      LastCheck[s] = -1;
    }
    LastChange.assign(Dims.size(), kFarChange);
    LowestHint.assign(Dims.size(), std::vector<double>());
    StepCount = 0;
    LaunchedStep = -1;
  }

  // one Lanczos step X_j -> (A_j, B_j, X_{j+1}) in every segment flagged in `active`
  void Step(const std::vector<unsigned char>& active) {
    LaunchProduct(active);
    FinishStep(active);
  }

  // first half of a step: enqueue W = H X_j for the segments flagged in `active` (nothing waits,
  // so the host can test convergence while the device multiplies)
  void LaunchProduct(const std::vector<unsigned char>& active) {
    const int j = StepCount;
    if (j + 1 > MaxBlocks) throw std::runtime_error("Krylov storage exhausted");
    Reserve(j + 1);
    Ops->Product(Block(j), Block(j + 1), active.data());
    Launched = active;
    LaunchedStep = j;
  }

  // second half: orthogonalise the product of the segments flagged in `active`, which must be
  // among those of the pending LaunchProduct
  void FinishStep(const std::vector<unsigned char>& active) {
    const int j = StepCount;
    if (LaunchedStep != j || Launched.size() != active.size())
      throw std::logic_error("FinishStep without the product of this step");
    std::vector<int> act;
    for (size_t s = 0; s < active.size(); ++s)
      if (active[s]) {
        if (!Launched[s]) throw std::logic_error("FinishStep on a segment whose product was not launched");
        act.push_back((int)s);
      }
    LaunchedStep = -1;
    double* W = Block(j + 1);
    std::vector<Rounds> r;
    Orthonormalize(W, j + 1, act, r, true);
    for (size_t a = 0; a < act.size(); ++a) {
      const int s = act[a];
      std::vector<T>& h = r[a].coef;
      for (int p = 0; p < K; ++p) {
        h[(size_t)p * K + p] = T(std::real(h[(size_t)p * K + p]));
        for (int q = p + 1; q < K; ++q) {
          T v = 0.5 * (h[(size_t)p * K + q] + S::Conj(h[(size_t)q * K + p]));
          h[(size_t)p * K + q] = v;
          h[(size_t)q * K + p] = S::Conj(v);
        }
      }
      A[s].push_back(h);
      B[s].push_back(r[a].R);
      DeadCols[s].push_back(Dead(r[a].good));
    }
    ++StepCount;
  }

  // the newest block of s has no live column: the Krylov space of s is complete
  bool Exhausted(int s) const {
    const std::vector<char>& d = DeadCols[s].back();
    return std::all_of(d.begin(), d.end(), [](char x) { return x != 0; });
  }

  // orthonormal Krylov vectors of s behind the projected matrix (its meaningful dimension)
  int64_t Live(int s) const {
    int64_t n = 0;
    for (size_t b = 0; b < A[s].size(); ++b)
      for (char x : DeadCols[s][b]) n += x ? 0 : 1;
    return n;
  }

  // eigenvalues of the projected matrix of s (dead columns removed); evecs: Ritz coefficients,
  // column-major over all m*k Krylov columns, the first Live(s) columns belong to `ev`
  void Eigen(int s, std::vector<double>& ev, std::vector<T>* evecs = nullptr) const {
    const int m = (int)A[s].size();
    std::vector<std::vector<T>> As = A[s];
    double bound = 0.0;  // Gershgorin bound of the matrix
    for (int b = 0; b < m; ++b)
      for (int p = 0; p < K; ++p) {
        double rs = 0.0;
        for (int q = 0; q < K; ++q) {
          rs += std::abs(A[s][b][(size_t)p * K + q]);
          if (b > 0) rs += std::abs(B[s][b - 1][(size_t)p * K + q]);
          if (b + 1 < m) rs += std::abs(B[s][b][(size_t)q * K + p]);
        }
        bound = std::max(bound, rs);
      }
    const double sentinel = 2.0 * bound + 1.0;
    for (int b = 0; b < m; ++b)
      for (int p = 0; p < K; ++p)
        if (DeadCols[s][b][p]) As[b][(size_t)p * K + p] = T(sentinel);
    BlockTridiagonalEigen(As, B[s], K, m, ev, evecs);
    ev.resize((size_t)Live(s));
  }

  // The min(count, Live(s)) lowest eigenvalues of s for block size 1: the projected matrix is
  // tridiagonal (a Hermitian one has the spectrum of the real one with |B_j| off the diagonal).
  void LowestEigen(int s, int count, std::vector<double>& ev) const {
    if (K != 1) {
      Eigen(s, ev);
      ev.resize(std::min<size_t>(ev.size(), (size_t)std::max(0, count)));
      return;
    }
    const int m = (int)A[s].size();
    std::vector<double> d(m), e(m > 0 ? m - 1 : 0);
    double bound = 0.0;
    for (int b = 0; b < m; ++b) {
      d[b] = std::real(A[s][b][0]);
      if (b + 1 < m) e[b] = std::abs(B[s][b][0]);
      bound = std::max(bound, std::fabs(d[b]) + (b > 0 ? std::fabs(e[b - 1]) : 0.0) + (b + 1 < m ? std::fabs(e[b]) : 0.0));
    }
    for (int b = 0; b < m; ++b)
      if (DeadCols[s][b][0]) d[b] = 2.0 * bound + 1.0;
    // per-segment entries only: the sector solves test their segments in parallel
    TridiagonalLowest(d, e, (int)std::min<int64_t>(count, Live(s)), ev, &LowestHint[s]);
    LowestHint[s] = ev;
  }

  // Block Lanczos convergence checks cost a band reduction; past the first ones they are spaced
  // so that at most ~10% extra blocks run after convergence.  The one-vector recurrence tests every
  // step once the wanted values moved by at most kNearChange of themselves at the last test, and
  // every fourth step before: a test (eleven bisections over the tridiagonal matrix) costs more than
  // a small sector's whole step, and a value still moving by 1e-8 of itself needs more than one
  // step to settle within the 1e-14 of the test, so at most three steps run past convergence.
  static constexpr double kNearChange = 1e-8, kFarChange = 1e300;
  bool CheckDue(int s, int nbrEigen) const {
    const int m = NbrBlocks(s);
    if (Exhausted(s)) return true;
    // a second test on the same matrix would compare the Ritz values with themselves
    if (LastCheck[s] == m) return false;
    if ((int64_t)m * K < nbrEigen + 2) return false;
    if (K == 1) return LastCheck[s] < 0 || LastChange[s] <= kNearChange || m - LastCheck[s] >= 4;
    return LastCheck[s] < 0 || m - LastCheck[s] >= std::max(1, m / 10);
  }

  // strong test of FullReorthogonalizedLanczosAlgorithm.cc:325: every wanted eigenvalue moved by
  // less than precision * |value| since the previous test of s (or the Krylov space is complete).
  // A value that is zero to within 3 MACHINE_PRECISION now and before counts as converged
  // (:335-338): exact zero modes otherwise move by rounding noise relative to themselves forever.
  static constexpr double kMachinePrecision = 1e-14;  // src/config.h:49
  bool TestConvergence(int s, int nbrEigen, double precision, std::vector<double>& current) {
    std::vector<double> ev;
    if (K == 1)
      LowestEigen(s, nbrEigen + 1, ev);
    else
      Eigen(s, ev);
    LastCheck[s] = NbrBlocks(s);
    const int want = std::min<int>(nbrEigen, (int)ev.size());
    current.assign(ev.begin(), ev.begin() + want);
    if (Exhausted(s)) {
      Previous[s] = current;
      return true;
    }
    bool ok = !Previous[s].empty() && (int)ev.size() > nbrEigen && (int)Previous[s].size() == want;
    // the largest relative move of a wanted value since the previous test (values within 3
    // MACHINE_PRECISION of zero now and before count as still), for the test cadence of CheckDue
    double change = kFarChange;
    if (!Previous[s].empty() && (int)Previous[s].size() == want) {
      change = 0.0;
      for (int e = 0; e < want; ++e)
        if (std::fabs(current[e]) > 3 * kMachinePrecision || std::fabs(Previous[s][e]) > 3 * kMachinePrecision)
          change = std::max(change, std::fabs(current[e] - Previous[s][e]) / std::max(std::fabs(current[e]), 1e-300));
    }
    LastChange[s] = change;
    if (ok)
      for (int e = 0; e < want; ++e)
        if (std::fabs(current[e] - Previous[s][e]) > precision * std::fabs(current[e]) &&
            (std::fabs(current[e]) > 3 * kMachinePrecision || std::fabs(Previous[s][e]) > 3 * kMachinePrecision)) {
          ok = false;
          break;
        }
    Previous[s] = current;
    return ok;
  }

  // Host half of a checkpoint: step count, seed of the next random replacement, and per segment
  // the projected blocks, dead columns and convergence history, in native (little-endian) byte
  // order.  The basis blocks X_0..X_Steps() are the caller's to store; restoring both continues
  // the recurrence exactly where it stopped.
  void SaveState(std::ostream& o) const {
    auto put = [&](const void* p, size_t n) { o.write(static_cast<const char*>(p), (std::streamsize)n); };
    const int32_t steps = StepCount, segs = (int32_t)Dims.size(), k = K;
    put(&steps, 4);
    put(&NextSeed, 8);
    put(&segs, 4);
    put(&k, 4);
    for (size_t s = 0; s < Dims.size(); ++s) {
      const int32_t m = (int32_t)A[s].size(), nd = (int32_t)DeadCols[s].size(), np = (int32_t)Previous[s].size();
      const int32_t last = LastCheck[s];
      put(&Dims[s], 8);
      put(&m, 4);
      put(&nd, 4);
      put(&np, 4);
      put(&last, 4);
      for (int b = 0; b < m; ++b) {
        put(A[s][b].data(), sizeof(T) * A[s][b].size());
        put(B[s][b].data(), sizeof(T) * B[s][b].size());
      }
      for (int b = 0; b < nd; ++b) put(DeadCols[s][b].data(), (size_t)K);
      put(Previous[s].data(), sizeof(double) * (size_t)np);
      put(&LastChange[s], sizeof(double));
    }
    if (!o) throw std::runtime_error("cannot write the Lanczos state");
  }

  void LoadState(std::istream& in) {
    auto get = [&](void* p, size_t n) {
      in.read(static_cast<char*>(p), (std::streamsize)n);
      if (!in) throw std::invalid_argument("truncated Lanczos state");
    };
    int32_t steps, segs, k;
    uint64_t seed;
    get(&steps, 4);
    get(&seed, 8);
    get(&segs, 4);
    get(&k, 4);
    if (segs != (int32_t)Dims.size() || k != K) throw std::invalid_argument("the Lanczos state has another block layout");
    if (steps < 0 || steps + 1 > MaxBlocks + 1) throw std::invalid_argument("the Lanczos state does not fit the Krylov storage");
    Reserve(steps);
    for (size_t s = 0; s < Dims.size(); ++s) {
      int64_t dim;
      int32_t m, nd, np, last;
      get(&dim, 8);
      get(&m, 4);
      get(&nd, 4);
      get(&np, 4);
      get(&last, 4);
      if (dim != Dims[s]) throw std::invalid_argument("Hamiltonian does not match dimension of stored vectors in resuming");
      if (m < 0 || m > steps || nd != m + 1 || np < 0 || np > (1 << 24) || last > m)
        throw std::invalid_argument("inconsistent Lanczos state");
      A[s].assign((size_t)m, std::vector<T>((size_t)K * K));
      B[s].assign((size_t)m, std::vector<T>((size_t)K * K));
      for (int b = 0; b < m; ++b) {
        get(A[s][b].data(), sizeof(T) * (size_t)K * K);
        get(B[s][b].data(), sizeof(T) * (size_t)K * K);
      }
      DeadCols[s].assign((size_t)nd, std::vector<char>((size_t)K));
      for (int b = 0; b < nd; ++b) get(DeadCols[s][b].data(), (size_t)K);
      Previous[s].assign((size_t)np, 0.0);
      get(Previous[s].data(), sizeof(double) * (size_t)np);
      get(&LastChange[s], sizeof(double));
      LastCheck[s] = last;
    }
    StepCount = steps;
    LaunchedStep = -1;
    NextSeed = seed;
  }

 private:
  // per segment: W_in = V_{0..nb} coef-part + Q R; coef = coefficients on the newest basis block
  struct Rounds {
    std::vector<T> coef, R;
    std::vector<char> good;
  };

  static std::vector<char> Dead(const std::vector<char>& good) {
    std::vector<char> d(good.size());
    for (size_t p = 0; p < good.size(); ++p) d[p] = good[p] ? 0 : 1;
    return d;
  }

  // Semidefinite Cholesky G = R^H R.  Column i is deficient when its pivot is below 1e-13 of its
  // squared norm gii (dependent on the previous columns) or when gii is below `collapse` times
  // ref[i], its squared norm before the projection that preceded (only rounding noise inside the
  // span of the basis survived).  Deficient columns get zero rows in R.
  static void CholeskySemidefinite(int k, const T* G, const double* ref, double collapse, std::vector<T>& R,
                                   std::vector<char>& good) {
    R.assign((size_t)k * k, T(0.0));
    good.assign(k, 0);
    for (int i = 0; i < k; ++i) {
      const double gii = std::real(G[(size_t)i * k + i]);
      double s = gii;
      for (int p = 0; p < i; ++p) s -= S::Norm(R[(size_t)p * k + i]);
      if (!(s > 0.0) || !(s > 1e-13 * gii) || !(gii > collapse * ref[i])) continue;
      good[i] = 1;
      const double rii = std::sqrt(s);
      R[(size_t)i * k + i] = T(rii);
      for (int j = i + 1; j < k; ++j) {
        T t = G[(size_t)i * k + j];
        for (int p = 0; p < i; ++p) t -= S::Conj(R[(size_t)p * k + i]) * R[(size_t)p * k + j];
        R[(size_t)i * k + j] = t / rii;
      }
    }
  }

  // inverse of R on its non-zero rows/columns (upper triangular), zero elsewhere
  static std::vector<T> PseudoInverse(int k, const std::vector<T>& R, const std::vector<char>& good) {
    std::vector<T> X((size_t)k * k, T(0.0));
    for (int j = 0; j < k; ++j) {
      if (!good[j]) continue;
      X[(size_t)j * k + j] = T(1.0) / R[(size_t)j * k + j];
      for (int i = j - 1; i >= 0; --i) {
        if (!good[i]) continue;
        T s(0.0);
        for (int p = i + 1; p <= j; ++p)
          if (good[p]) s += R[(size_t)i * k + p] * X[(size_t)p * k + j];
        X[(size_t)i * k + j] = -s / R[(size_t)i * k + i];
      }
    }
    return X;
  }

  static std::vector<T> Mul(int k, const std::vector<T>& X, const std::vector<T>& Y) {
    std::vector<T> Z((size_t)k * k, T(0.0));
    for (int i = 0; i < k; ++i)
      for (int p = 0; p < k; ++p)
        for (int j = 0; j < k; ++j) Z[(size_t)i * k + j] += X[(size_t)i * k + p] * Y[(size_t)p * k + j];
    return Z;
  }

  // the basis blocks [b0, b0 + nb) in pieces that lie in one storage chunk each: f(ptr, first, count)
  template <class F>
  void ForPieces(int b0, int nb, F f) const {
    for (const Chunk& c : Chunks) {
      const int lo = std::max(b0, c.first), hi = std::min(b0 + nb, c.first + c.count);
      if (lo < hi) f(c.ptr + (int64_t)(lo - c.first) * BlockDoubles, lo - b0, hi - lo);
    }
  }
  bool OnePiece(int b0, int nb) const {
    for (const Chunk& c : Chunks)
      if (b0 >= c.first && b0 + nb <= c.first + c.count) return true;
    return false;
  }
  // out[a][b][KR x KR] = V_(b0 + b)^T W over segment a of `mask` (na of them), b < nb
  void BasisDots(int b0, int nb, const double* W, const unsigned char* mask, size_t na, double* out) {
    if (OnePiece(b0, nb)) {
      Ops->Dots(Block(b0), nb, W, mask, out);
      return;
    }
    const size_t KK = (size_t)KR * KR;
    ForPieces(b0, nb, [&](const double* V, int off, int cnt) {
      std::vector<double> part(na * cnt * KK);
      Ops->Dots(V, cnt, W, mask, part.data());
      for (size_t a = 0; a < na; ++a)
        std::copy(part.begin() + a * cnt * KK, part.begin() + (a + 1) * cnt * KK,
                  out + (a * nb + off) * KK);
    });
  }
  // W += sum_b V_(b0 + b) C[a][b] over segment a
  void BasisAxpy(int b0, int nb, double* W, const unsigned char* mask, size_t na, const double* C) {
    const size_t KK = (size_t)KR * KR;
    ForPieces(b0, nb, [&](const double* V, int off, int cnt) {
      std::vector<double> part(na * cnt * KK);
      for (size_t a = 0; a < na; ++a)
        std::copy(C + (a * nb + off) * KK, C + (a * nb + off + cnt) * KK, part.begin() + a * cnt * KK);
      Ops->Axpy(V, cnt, W, mask, part.data());
    });
  }
  // W += sum_b V_(b0 + b) C[a][b] over segment a, then gram[a] = W^T W
  void BasisUpdate(int b0, int nb, double* W, const unsigned char* mask, size_t na, const double* C, double* gram) {
    if (OnePiece(b0, nb)) {
      Ops->Update(Block(b0), nb, W, mask, C, nullptr, nullptr, gram);
      return;
    }
    BasisAxpy(b0, nb, W, mask, na, C);
    Ops->Dots(W, 1, W, mask, gram);
  }

  std::vector<unsigned char> Mask(const std::vector<int>& act, const std::vector<int>& idx) const {
    std::vector<unsigned char> mask(Dims.size(), 0);
    for (int i : idx) mask[act[i]] = 1;
    return mask;
  }

  // Coefficients of one projection from the device-layout dots C (per segment of idx, nb blocks of
  // KR x KR): the coefficient on the newest block goes to out[i].coef (times out[i].R after the
  // first projection, see OrthoRounds) and the squared norms removed to ref[i]; C is negated in
  // place, ready to be added to W.
  // CT: the coefficients as k x k scalar blocks (segment a, block b at (a * nb + b) * k * k); C
  // receives their negation in the device layout of the axpy.
  void Coefficients(const std::vector<T>& CTall, double* C, int nb, const std::vector<int>& idx,
                    std::vector<Rounds>& out, bool first, std::vector<std::vector<double>>& ref) const {
    const int kk = K * K, KK = KR * KR;
    const size_t per = (size_t)nb * KK;
    std::vector<T> CT((size_t)kk);
    for (size_t a = 0; a < idx.size(); ++a) {
      Rounds& o = out[idx[a]];
      for (int b = 0; b < nb; ++b) {
        double* blk = &C[a * per + (size_t)b * KK];
        std::copy(CTall.begin() + (a * nb + b) * kk, CTall.begin() + (a * nb + b + 1) * kk, CT.begin());
        for (int p = 0; p < K; ++p)
          for (int q = 0; q < K; ++q) ref[idx[a]][q] += S::Norm(CT[(size_t)p * K + q]);
        if (b == nb - 1) {
          std::vector<T> add = first ? CT : Mul(K, CT, o.R);
          for (int q = 0; q < kk; ++q) o.coef[q] += add[q];
        }
        for (auto& v : CT) v = -v;
        S::ToDevice(K, CT.data(), blk);
      }
    }
  }

  // Cholesky-QR factor of the Gram blocks G of the segments of idx, with the deficiency tests of
  // CholeskySemidefinite (ref[i]: squared norms removed by the projection just before, if any):
  // out[i].R <- Rr out[i].R, out[i].good, the device-layout inverse of Rr into M and, if asked,
  // its T form into inv.
  void Factor(const double* G, const std::vector<int>& idx, std::vector<Rounds>& out,
              std::vector<std::vector<double>>& ref, double collapse, double* M,
              std::vector<std::vector<T>>* inv) const {
    const int kk = K * K, KK = KR * KR;
    std::vector<T> GT((size_t)kk);
    if (inv) inv->assign(idx.size(), std::vector<T>());
    for (size_t a = 0; a < idx.size(); ++a) {
      Rounds& o = out[idx[a]];
      std::vector<double>& r = ref[idx[a]];
      S::FromGram(K, &G[a * KK], GT.data());
      for (int p = 0; p < K; ++p) r[p] += std::real(GT[(size_t)p * K + p]);
      std::vector<T> Rr;
      std::vector<char> good;
      CholeskySemidefinite(K, GT.data(), r.data(), collapse, Rr, good);
      std::vector<T> X = PseudoInverse(K, Rr, good);
      S::ToDevice(K, X.data(), &M[a * KK]);
      if (inv) (*inv)[a] = X;
      o.R = Mul(K, Rr, o.R);
      o.good = good;
    }
  }

  // One Cholesky-QR of W over the segments act[i], i in idx (the start block, before any basis).
  void CholQR(double* W, const std::vector<int>& act, const std::vector<int>& idx, std::vector<Rounds>& out,
              std::vector<std::vector<double>>& ref, double collapse) {
    const int KK = KR * KR;
    std::vector<unsigned char> mask = Mask(act, idx);
    std::vector<double> G(idx.size() * KK), M(idx.size() * KK);
    Ops->Dots(W, 1, W, mask.data(), G.data());
    Factor(G.data(), idx, out, ref, collapse, M.data(), nullptr);
    Ops->Rmul(W, mask.data(), M.data());
  }

  // Two rounds of projection + Cholesky-QR of W over the segments in `act`: W_in = V C + Q R
  // with C = C1 + C2 R1 and R = R2 R1.  Deficient columns of Q are zero.
  //
  // For the new block of a step (W = H X_j, `local`), the first round projects on the two newest
  // blocks only: in exact arithmetic W has no component on older blocks (the block three-term
  // recurrence), so this round removes the large coefficients A_j and B_j^H, as the local
  // recurrence of FullReorthogonalizedLanczosAlgorithm.cc:238-245 does.  The second round projects
  // on the whole basis (the full reorthogonalisation of :246-258) and removes the rounding left on
  // every block.  Both rounds always run: one projection leaves the basis orthogonal to ~1e-13
  // only, and the Ritz values then move by more than the 1e-14 relative change that the
  // convergence test asks for.  Random replacement vectors have components on every block, so both
  // of their rounds are full (`local` false).
  //
  // Passes over the basis V: the dots C = V^T W, then the update W - V C (with the Gram matrix of
  // the result), then W R^-1.  With the local first round a step reads the whole basis twice and
  // its two newest blocks twice more.
  void OrthoRounds(double* W, int nb, const std::vector<int>& act, std::vector<Rounds>& out, bool local = false) {
    const int kk = K * K, KK = KR * KR;
    const size_t na = act.size();
    out.assign(na, Rounds());
    for (auto& o : out) {
      o.coef.assign((size_t)kk, T(0.0));
      o.R.assign((size_t)kk, T(0.0));
      for (int p = 0; p < K; ++p) o.R[(size_t)p * K + p] = T(1.0);
      o.good.assign(K, 1);
    }
    std::vector<int> all(na);
    for (size_t a = 0; a < na; ++a) all[a] = (int)a;
    std::vector<std::vector<double>> ref(na, std::vector<double>(K, 0.0));
    if (nb == 0) {
      // noise left by one Cholesky-QR is ~1e-15 of the column
      CholQR(W, act, all, out, ref, 1e-24);
      for (auto& r : ref) std::fill(r.begin(), r.end(), 0.0);
      CholQR(W, act, all, out, ref, 1e-16);
      return;
    }
    std::vector<unsigned char> mask = Mask(act, all);
    std::vector<double> C(na * (size_t)nb * KK), G(na * KK), M1(na * KK), M2(na * KK);
    std::vector<T> CT(na * (size_t)nb * kk);
    // scalar blocks of the dots over nv blocks, in the device layout
    auto fromDevice = [&](const double* src, int nv) {
      const size_t per = (size_t)nv * KK;
      for (size_t a = 0; a < na; ++a)
        for (int b = 0; b < nv; ++b) S::FromGram(K, &src[a * per + (size_t)b * KK], &CT[(a * nv + b) * kk]);
    };
    // first round: the two newest blocks (local) or all of them; the newest is the last either way
    const int nb1 = local ? std::min(nb, 2) : nb;
    BasisDots(nb - nb1, nb1, W, mask.data(), na, C.data());
    fromDevice(C.data(), nb1);
    Coefficients(CT, C.data(), nb1, all, out, true, ref);
    if (local && K == 1 && W == Block(nb)) {
      // A step's new vector follows the basis in the storage (W = X_nb), so the dots of the second
      // round and the squared norm of the first round's result come from one pass over blocks
      // 0 .. nb (the last against itself): one synchronisation less per step.  The dots are those
      // of W before its normalisation W / |W|; the coefficients follow as (V^H W) / |W|.  Blocks
      // of k > 1 keep the separate passes: (V^H W) R1^-1 loses accuracy with the conditioning of
      // R1 when the block's columns are nearly dependent, which the deficiency tests must see.
      BasisAxpy(nb - nb1, nb1, W, mask.data(), na, C.data());
      std::vector<double> Cx(na * (size_t)(nb + 1) * KK);
      BasisDots(0, nb + 1, W, mask.data(), na, Cx.data());
      for (size_t a = 0; a < na; ++a) {
        const double* src = &Cx[a * (size_t)(nb + 1) * KK];
        std::copy(src, src + (size_t)nb * KK, &C[a * (size_t)nb * KK]);
        std::copy(src + (size_t)nb * KK, src + (size_t)(nb + 1) * KK, &G[a * KK]);
      }
      std::vector<std::vector<T>> inv1;
      Factor(G.data(), all, out, ref, 1e-24, M1.data(), &inv1);
      for (auto& r : ref) std::fill(r.begin(), r.end(), 0.0);
      Ops->Rmul(W, mask.data(), M1.data());
      fromDevice(C.data(), nb);
      std::vector<T> blk((size_t)kk);
      for (size_t a = 0; a < na; ++a)
        for (int b = 0; b < nb; ++b) {
          T* c = &CT[(a * nb + b) * kk];
          std::copy(c, c + kk, blk.begin());
          const std::vector<T> scaled = Mul(K, blk, inv1[a]);
          std::copy(scaled.begin(), scaled.end(), c);
        }
      Coefficients(CT, C.data(), nb, all, out, false, ref);
      BasisUpdate(0, nb, W, mask.data(), na, C.data(), G.data());
      Factor(G.data(), all, out, ref, 1e-16, M2.data(), nullptr);
      Ops->Rmul(W, mask.data(), M2.data());
      return;
    }
    BasisUpdate(nb - nb1, nb1, W, mask.data(), na, C.data(), G.data());
    // noise left by one projection is ~1e-15 of the column
    Factor(G.data(), all, out, ref, 1e-24, M1.data(), nullptr);
    for (auto& r : ref) std::fill(r.begin(), r.end(), 0.0);
    Ops->Rmul(W, mask.data(), M1.data());
    BasisDots(0, nb, W, mask.data(), na, C.data());
    fromDevice(C.data(), nb);
    Coefficients(CT, C.data(), nb, all, out, false, ref);
    BasisUpdate(0, nb, W, mask.data(), na, C.data(), G.data());
    // after the second projection, a column that kept less than 1e-8 of itself was
    // (normalised) noise inside the span of the basis
    Factor(G.data(), all, out, ref, 1e-16, M2.data(), nullptr);
    Ops->Rmul(W, mask.data(), M2.data());
  }

  // W (nb basis blocks before it) -> orthonormal Q with W_in = V C + Q R per segment of `act`;
  // deficient columns are replaced by random directions where the segment has room left
  void Orthonormalize(double* W, int nb, const std::vector<int>& act, std::vector<Rounds>& out, bool local = false) {
    const int kk = K * K, KK = KR * KR;
    OrthoRounds(W, nb, act, out, local);
    std::vector<int> need, needSeg;
    for (size_t a = 0; a < act.size(); ++a)
      if (std::find(out[a].good.begin(), out[a].good.end(), 0) != out[a].good.end()) {
        need.push_back((int)a);
        needSeg.push_back(act[a]);
      }
    if (need.empty()) return;
    // W = Q (deficient columns zero) + random vectors in the deficient columns
    std::vector<unsigned char> mask(Dims.size(), 0);
    for (int s : needSeg) mask[s] = 1;
    double* Rnd = Scratch();
    Ops->Random(Rnd, NextSeed++);
    std::vector<double> Sel(need.size() * KK, 0.0);
    for (size_t i = 0; i < need.size(); ++i) {
      std::vector<T> E((size_t)kk, T(0.0));
      for (int p = 0; p < K; ++p)
        if (!out[need[i]].good[p]) E[(size_t)p * K + p] = T(1.0);
      S::ToDevice(K, E.data(), &Sel[i * KK]);
    }
    Ops->Axpy(Rnd, 1, W, mask.data(), Sel.data());
    // Q_rep = V C' + Q' R'  =>  W_in = V (C + C' R) + Q' (R' R)  (R has zero rows where Q_rep is random)
    std::vector<Rounds> rep;
    OrthoRounds(W, nb, needSeg, rep);
    for (size_t i = 0; i < need.size(); ++i) {
      Rounds& o = out[need[i]];
      std::vector<T> add = Mul(K, rep[i].coef, o.R);
      for (int q = 0; q < kk; ++q) o.coef[q] += add[q];
      o.R = Mul(K, rep[i].R, o.R);
      o.good = rep[i].good;
    }
  }

  BlockOps* Ops;
  std::vector<int64_t> Dims;
  int K, KR, MaxBlocks;
  int Capacity;  // Krylov blocks the storage holds (blocks 0 .. Capacity, then the scratch block)
  GrowFn Grow;
  struct Chunk {
    double* ptr;
    int first, count;  // Krylov blocks [first, first + count)
  };
  std::vector<Chunk> Chunks;
  double* ScratchBlock;
  int64_t BlockDoubles;
  int StepCount = 0;
  int LaunchedStep = -1;               // step whose product LaunchProduct enqueued (-1: none)
  std::vector<unsigned char> Launched;  // segments of that product
  uint64_t NextSeed = 1;
  std::vector<std::vector<std::vector<T>>> A, B;
  std::vector<std::vector<std::vector<char>>> DeadCols;  // per segment, per block X_b, per column
  std::vector<std::vector<double>> Previous;
  std::vector<int> LastCheck;
  std::vector<double> LastChange;  // per segment: the largest relative move at the last test
  mutable std::vector<std::vector<double>> LowestHint;  // per segment: the last LowestEigen result
};

}  // namespace dhh
