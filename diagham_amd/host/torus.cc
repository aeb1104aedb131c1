// Fermions on the torus in one Ky sector (FermionOnTorus with a momentum constraint,
// FQHE/src/HilbertSpace/FermionOnTorus.cc:106) and their two-body Hamiltonians
// (ParticleOnTorusGenericHamiltonian, ParticleOnTorusCoulombHamiltonian).
//
// The basis is the momentum-space basis with Nx = 1 (orbital ky, states of total momentum Ky mod
// Nphi in descending order; AA / AdAd of FermionOnTorus are those of FermionOnSphere), so the
// device side is a momentum sector set and its gather kernel.  The factors are real; they are
// handed over in that set's complex layout with zero imaginary parts.
#include <cmath>
#include <stdexcept>
#include <vector>

#include "diagham_host.hpp"

namespace dhh {

namespace {

// monomial coefficients of the Laguerre polynomial L_n (LaguerrePolynomial(n) of the reference):
// c_k = (-1)^k C(n, k) / k!
std::vector<double> LaguerreCoefficients(int n) {
  std::vector<double> c(n + 1);
  double binom = 1.0, fact = 1.0;
  for (int k = 0; k <= n; ++k) {
    if (k > 0) {
      binom *= (double)(n - k + 1) / (double)k;
      fact *= (double)k;
    }
    c[k] = ((k & 1) ? -binom : binom) / fact;
  }
  return c;
}

double Horner(const std::vector<double>& c, double x) {
  double r = 0.0;
  for (int k = (int)c.size() - 1; k >= 0; --k) r = r * x + c[k];
  return r;
}

// V(q) of one torus Hamiltonian as a function of Q2 = ratio N2^2 + N1^2 / ratio (integer
// reciprocal coordinates N1, N2); zero q is handled by the callers
struct TorusInteraction {
  int NbrFlux;
  double Ratio, InvRatio, PiOnM;
  int CoulombLevel;  // < 0: pseudopotential form
  std::vector<double> Pseudo;
  std::vector<std::vector<double>> Laguerre;  // per pseudopotential index
  std::vector<double> FormFactor;             // Coulomb: L_n

  TorusInteraction(int nbrFlux, double ratio, const std::vector<double>& pseudo, int coulombLevel)
      : NbrFlux(nbrFlux), Ratio(ratio), InvRatio(1.0 / ratio), PiOnM(M_PI / nbrFlux), CoulombLevel(coulombLevel),
        Pseudo(pseudo) {
    for (size_t i = 0; i < Pseudo.size(); ++i) Laguerre.push_back(LaguerreCoefficients((int)i));
    if (CoulombLevel >= 0) FormFactor = LaguerreCoefficients(CoulombLevel);
  }

  // sum_i V_i L_i(2 pi Q2 / Nphi) (ParticleOnTorusGenericHamiltonian.cc:318)
  double PseudoSum(double x) const {
    double t = 0.0;
    for (size_t i = 0; i < Pseudo.size(); ++i)
      if (Pseudo[i] != 0.0) t += Pseudo[i] * Horner(Laguerre[i], x);
    return t;
  }

  // GetVofQ (ParticleOnTorusCoulombHamiltonian.cc:284) at half of q^2
  double CoulombVofQ(double q2Half) const {
    if (q2Half == 0.0) return 0.0;
    const double ff = Horner(FormFactor, q2Half);
    return ff * ff / std::sqrt(2.0 * q2Half) * std::exp(-q2Half);
  }

  // the first term of one N2 column and its "precision" seed, then the N1 > 0 terms
  double Column0(double n2, double& precision) const {
    const double q2 = Ratio * n2 * n2;
    if (CoulombLevel >= 0) {
      const double v = CoulombVofQ(PiOnM * q2);
      precision = (n2 != 0.0) ? v : 1.0;
      return v;
    }
    if (n2 != 0.0) {
      const double v = std::exp(-PiOnM * q2) * PseudoSum(2.0 * PiOnM * q2);
      precision = v;
      return v;
    }
    precision = 1.0;
    return PseudoSum(0.0);
  }
  double ColumnTerm(double n1, double n2) const {
    const double q2 = InvRatio * n1 * n1 + Ratio * n2 * n2;
    if (CoulombLevel >= 0) return 2.0 * CoulombVofQ(PiOnM * q2);
    return 2.0 * std::exp(-PiOnM * q2) * PseudoSum(2.0 * PiOnM * q2);
  }

  // EvaluateInteractionCoefficient (ParticleOnTorusGenericHamiltonian.cc:318,
  // ParticleOnTorusCoulombHamiltonian.cc:138): N2 = m1 - m4 + j Nphi, outwards in both directions,
  // each column summed over N1 with cos(2 pi N1 (m1 - m3) / Nphi), until terms stop changing the sums
  double Coefficient(int m1, int m3, int m4) const {
    const double cosArg = -((double)(m1 - m3)) * PiOnM * 2.0;
    double sum = 0.0;
    for (int dir = 0; dir < 2; ++dir) {
      double n2 = (double)(m1 - m4 - (dir ? NbrFlux : 0));
      double column = (dir == 0 || CoulombLevel >= 0) ? 1.0 : sum;
      while (std::fabs(sum) + std::fabs(column) != std::fabs(sum)) {
        double precision;
        column = Column0(n2, precision);
        for (double n1 = 1.0; std::fabs(column) + precision != std::fabs(column); n1 += 1.0) {
          precision = ColumnTerm(n1, n2);
          column += precision * std::cos(n1 * cosArg);
        }
        sum += column;
        n2 += dir ? -NbrFlux : NbrFlux;
      }
    }
    return CoulombLevel >= 0 ? sum / (2.0 * NbrFlux) : sum / NbrFlux;
  }
};

}  // namespace

void TorusInteractionFactors(int nbrFlux, double ratio, const std::vector<double>& pseudo, int coulombLevel,
                             std::vector<int>& sumPairCount, std::vector<int>& sumPairs,
                             std::vector<double>& factors) {
  if (nbrFlux < 2 || nbrFlux > 62) throw std::invalid_argument("the torus needs 2 to 62 flux quanta");
  if (!(ratio > 0.0)) throw std::invalid_argument("the aspect ratio must be positive");
  if (coulombLevel < 0 && pseudo.empty()) throw std::invalid_argument("no pseudopotentials and no Coulomb level");
  const TorusInteraction v(nbrFlux, ratio, pseudo, coulombLevel);
  // AbstractQHEOnTorusHamiltonian::GetIndices (:80): pairs m1 < m2 by (m1 + m2) mod Nphi
  std::vector<std::vector<int>> lists(nbrFlux);
  for (int m1 = 0; m1 < nbrFlux - 1; ++m1)
    for (int m2 = m1 + 1; m2 < nbrFlux; ++m2) {
      lists[(m1 + m2) % nbrFlux].push_back(m1);
      lists[(m1 + m2) % nbrFlux].push_back(m2);
    }
  sumPairCount.assign(nbrFlux, 0);
  sumPairs.clear();
  factors.clear();
  for (int s = 0; s < nbrFlux; ++s) {
    const int n = (int)lists[s].size() / 2;
    sumPairCount[s] = n;
    sumPairs.insert(sumPairs.end(), lists[s].begin(), lists[s].end());
    // ParticleOnTorusGenericHamiltonian::EvaluateInteractionFactors (:244), fermions: the
    // antisymmetrised combination of the four orderings; F[annihilated j1][created j2]
    for (int j1 = 0; j1 < n; ++j1) {
      const int m1 = lists[s][2 * j1], m2 = lists[s][2 * j1 + 1];
      for (int j2 = 0; j2 < n; ++j2) {
        const int m3 = lists[s][2 * j2], m4 = lists[s][2 * j2 + 1];
        const double f = v.Coefficient(m1, m3, m4) + v.Coefficient(m2, m4, m3) - v.Coefficient(m1, m4, m3) -
                         v.Coefficient(m2, m3, m4);
        factors.push_back(f);
        factors.push_back(0.0);
      }
    }
  }
}

}  // namespace dhh
