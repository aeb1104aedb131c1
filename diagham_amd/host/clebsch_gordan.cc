// Clebsch-Gordan coefficients and LLL Coulomb pseudopotentials for the host side.
#include <quadmath.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>

#include "diagham_host.hpp"

namespace dhh {

namespace {
// n! in 113-bit precision, each entry the running product of the previous one (the same operations,
// and so the same bits, as multiplying 2 .. n in turn); tabulated below kFactTable
constexpr int kFactTable = 1700;  // 1700! < 1e4800, inside the __float128 range
const std::vector<__float128>& FactTable() {
  static const std::vector<__float128> table = [] {
    std::vector<__float128> t(kFactTable);
    t[0] = 1;
    if (kFactTable > 1) t[1] = 1;
    for (int i = 2; i < kFactTable; ++i) t[i] = t[i - 1] * i;
    return t;
  }();
  return table;
}

__float128 fact128(int n) {
  if (n < kFactTable) return FactTable()[n < 0 ? 0 : n];
  __float128 r = 1;
  for (int i = 2; i <= n; ++i) r *= i;
  return r;
}

// Racah's formula; all arguments doubled (j1, m1, j2, m2, J), M = m1 + m2.
double racah(int j1, int m1, int j2, int m2, int J) {
  int M = m1 + m2;
  if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(M) > J) return 0.0;
  if (((j1 + m1) & 1) || ((j2 + m2) & 1) || ((J + M) & 1)) return 0.0;
  if (J < std::abs(j1 - j2) || J > j1 + j2 || ((j1 + j2 + J) & 1)) return 0.0;
  int a = (J + j1 - j2) / 2, b = (J - j1 + j2) / 2, c = (j1 + j2 - J) / 2, d = (j1 + j2 + J) / 2 + 1;
  __float128 pre = (__float128)(J + 1) * fact128(a) * fact128(b) * fact128(c) / fact128(d);
  pre *= fact128((J + M) / 2) * fact128((J - M) / 2) * fact128((j1 - m1) / 2) * fact128((j1 + m1) / 2) *
         fact128((j2 - m2) / 2) * fact128((j2 + m2) / 2);
  int kmin = std::max(0, std::max((j2 - J - m1) / 2, (j1 - J + m2) / 2));
  int kmax = std::min(c, std::min((j1 - m1) / 2, (j2 + m2) / 2));
  __float128 sum = 0;
  for (int k = kmin; k <= kmax; ++k) {
    __float128 den = fact128(k) * fact128(c - k) * fact128((j1 - m1) / 2 - k) * fact128((j2 + m2) / 2 - k) *
                     fact128((J - j2 + m1) / 2 + k) * fact128((J - j1 - m2) / 2 + k);
    sum += ((k & 1) ? -1 : 1) / den;
  }
  return (double)(sqrtq(pre) * sum);
}

__float128 binom128(int n, int k) {
  if (k < 0 || k > n) return 0;
  __float128 r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}
}  // namespace

// The table of (j1, j2) is computed once per process and shared: every Hamiltonian of a sector
// loop over one sphere builds the same one.
ClebschGordanCoefficients::ClebschGordanCoefficients(int j1, int j2) : J1(j1), J2(j2) {
  static std::mutex mutex;
  static std::map<std::pair<int, int>, std::shared_ptr<const std::vector<double>>> cache;
  {
    std::lock_guard<std::mutex> lock(mutex);
    auto it = cache.find({j1, j2});
    if (it != cache.end()) {
      table = it->second;
      return;
    }
  }
  int nj = j1 + j2 + 1;
  auto t = std::make_shared<std::vector<double>>((size_t)(j1 + 1) * (j2 + 1) * nj, 0.0);
  for (int p1 = 0; p1 <= j1; ++p1)
    for (int p2 = 0; p2 <= j2; ++p2) {
      int m1 = 2 * p1 - j1, m2 = 2 * p2 - j2;
      for (int j = std::max(std::abs(j1 - j2), std::abs(m1 + m2)); j <= j1 + j2; j += 2)
        (*t)[((size_t)p1 * (j2 + 1) + p2) * nj + j] = racah(j1, m1, j2, m2, j);
    }
  std::lock_guard<std::mutex> lock(mutex);
  table = cache.emplace(std::make_pair(j1, j2), std::move(t)).first->second;
}

double ClebschGordanCoefficients::Coefficient(int j1, int j2, int m1, int m2, int j) {
  return racah(j1, m1, j2, m2, j);
}

double ClebschGordanCoefficients::GetCoefficient(int m1, int m2, int j) const {
  if (m1 < -J1 || m1 > J1 || m2 < -J2 || m2 > J2 || j < 0 || j > J1 + J2) return 0.0;
  int p1 = (J1 + m1) >> 1, p2 = (J2 + m2) >> 1;
  return (*table)[((size_t)p1 * (J2 + 1) + p2) * (J1 + J2 + 1) + j];
}

std::vector<double> EvaluatePseudopotentials(int nbrFlux) {
  std::vector<double> v(nbrFlux + 1, 0.0);
  const int lz = nbrFlux;  // twice S
  __float128 norm = binom128(2 * lz + 2, lz + 1);
  norm *= norm;
  for (int L = 0; L <= lz; ++L) {
    __float128 t = binom128(2 * lz - 2 * L, lz - L) * binom128(2 * lz + 2 * L + 2, lz + L + 1) / norm;
    v[lz - L] = (double)(2 * t / sqrtq((__float128)lz / 2));
  }
  return v;
}

std::vector<double> RoundToFileDigits(const std::vector<double>& v) {
  std::vector<double> r(v.size());
  char buf[64];
  for (size_t i = 0; i < v.size(); ++i) {
    std::snprintf(buf, sizeof(buf), "%.14g", v[i]);
    r[i] = std::strtod(buf, nullptr);
  }
  return r;
}

bool FQHESphereGetPseudopotentials(const std::string& fileName, int lzMax, std::vector<double>& pseudo) {
  std::ifstream f(fileName);
  if (!f) return false;
  std::string line;
  pseudo.assign(lzMax + 1, 0.0);
  bool found = false;
  while (std::getline(f, line)) {
    size_t hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    std::string key = line.substr(0, eq);
    key.erase(0, key.find_first_not_of(" \t"));
    key.erase(key.find_last_not_of(" \t") + 1);
    if (key != "Pseudopotentials") continue;
    std::istringstream vals(line.substr(eq + 1));
    double x;
    int i = 0;
    while (vals >> x) {
      if (i <= lzMax) pseudo[i] = x;
      ++i;
    }
    found = true;
  }
  return found;
}

}  // namespace dhh
