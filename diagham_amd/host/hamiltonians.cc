// Interaction factors of the generic two-body Hamiltonian on the sphere and the spin chain.
#include <quadmath.h>

#include <cmath>
#include <stdexcept>

#include "diagham_host.hpp"

namespace dhh {

// ParticleOnSphereGenericHamiltonian::EvaluateInteractionFactors (fermions,
// ParticleOnSphereGenericHamiltonian.cc:493-560):
//   F(m1 m2 -> m3 m4) = -4 sum_L V_{lzMax-L} <m1 m2|L> <m3 m4|L>
// over the pair angular momenta L whose relative momentum lzMax - L is odd, for every
// annihilated pair m1 > m2 and created pair m3 > m4 = m1 + m2 - m3 (orbital indices,
// l_z = index - lzMax/2).  Accumulated in 113-bit precision, then rounded.
void ParticleOnSphereGenericHamiltonian::EvaluateInteractionFactors(int lzMax, const std::vector<double>& pseudo,
                                                                     std::vector<double>& factors,
                                                                     std::vector<int>& pairOffset) {
  const int nOrb = lzMax + 1;
  ClebschGordanCoefficients cg(lzMax, lzMax);
  const int parity = (lzMax & 1) ? 0 : 1;  // L & 1 == parity  <=>  lzMax - L odd
  factors.clear();
  pairOffset.assign((size_t)nOrb * (nOrb - 1) / 2 + 1, 0);
  for (int m1 = 0; m1 < nOrb; ++m1)
    for (int m2 = 0; m2 < m1; ++m2) {
      pairOffset[(size_t)m1 * (m1 - 1) / 2 + m2] = (int)factors.size();
      int sum = m1 + m2;
      int m3lo = sum / 2 + 1, m3hi = std::min(sum, lzMax);
      int d1 = 2 * m1 - lzMax, d2 = 2 * m2 - lzMax;
      for (int m3 = m3lo; m3 <= m3hi; ++m3) {
        int m4 = sum - m3;
        int d3 = 2 * m3 - lzMax, d4 = 2 * m4 - lzMax;
        __float128 acc = 0;
        for (int J = std::abs(d1 + d2); J <= 2 * lzMax; J += 2) {
          int L = J >> 1;
          if ((L & 1) != parity) continue;
          acc += (__float128)pseudo[lzMax - L] * (__float128)cg.GetCoefficient(d1, d2, J) *
                 (__float128)cg.GetCoefficient(d3, d4, J);
        }
        factors.push_back((double)(-4 * acc));
      }
    }
  pairOffset.back() = (int)factors.size();
}

// Total angular momentum L^2 = L^- L^+ + Lz^2 + Lz of N fermions on the sphere, in the two-body
// factor layout above (the operator of ParticleOnSphereL2Hamiltonian,
// FQHE/src/Hamiltonian/ParticleOnSphereL2Hamiltonian.cc:240, derived here from
// L^+ = sum_p c_p a+_{p+1} a_p with c_p = sqrt((lzMax - p)(p + 1))):
//   L^- L^+ = sum_p c_p^2 n_p - sum_{p,q} c_p c_q a+_p a+_{q+1} a_{p+1} a_q.
// The two-body part is reordered to a+_m3 a+_m4 a_n1 a_n2 (n1 > n2, m3 > m4); the one-body part,
// inside a sector of N particles, is sum_{n1 > n2} (w_n1 + w_n2) n_n1 n_n2 / (N - 1) with
// n_n1 n_n2 = -a+_n1 a+_n2 a_n1 a_n2.  Everything is scaled by l2Factor; Lz^2 + Lz is a constant
// of the sector, returned as the shift l2Factor * Lz (Lz + 1), Lz = totalLz / 2.
void ParticleOnSphereL2Hamiltonian::EvaluateInteractionFactors(int lzMax, int nbrParticles, int totalLz,
                                                                double l2Factor, std::vector<double>& factors,
                                                                std::vector<int>& pairOffset, double& shift) {
  if (nbrParticles < 2) throw std::invalid_argument("L^2 as a two-body term needs at least two particles");
  const int nOrb = lzMax + 1;
  // zero table in the layout: pair (m1 > m2), created m3 from (m1 + m2) / 2 + 1 up
  factors.clear();
  pairOffset.assign((size_t)nOrb * (nOrb - 1) / 2 + 1, 0);
  for (int m1 = 0; m1 < nOrb; ++m1)
    for (int m2 = 0; m2 < m1; ++m2) {
      pairOffset[(size_t)m1 * (m1 - 1) / 2 + m2] = (int)factors.size();
      const int sum = m1 + m2;
      factors.resize(factors.size() + (std::min(sum, lzMax) - (sum / 2 + 1) + 1), 0.0);
    }
  pairOffset.back() = (int)factors.size();
  auto at = [&](int n1, int n2, int m3) -> double& {
    return factors[pairOffset[(size_t)n1 * (n1 - 1) / 2 + n2] + m3 - ((n1 + n2) / 2 + 1)];
  };
  std::vector<double> c(nOrb);
  for (int p = 0; p < nOrb; ++p) c[p] = std::sqrt((double)(lzMax - p) * (p + 1));
  for (int p = 0; p < lzMax; ++p)
    for (int q = 0; q < lzMax; ++q) {
      if (p + 1 == q || p == q + 1) continue;  // a_{p+1} a_q or a+_p a+_{q+1} vanishes
      const int a0 = p + 1, a1 = q, b0 = p, b1 = q + 1;
      const int n1 = std::max(a0, a1), n2 = std::min(a0, a1), m3 = std::max(b0, b1);
      const double sign = (a0 > a1 ? 1.0 : -1.0) * (b0 > b1 ? 1.0 : -1.0);
      at(n1, n2, m3) -= l2Factor * sign * c[p] * c[q];
    }
  for (int n1 = 0; n1 < nOrb; ++n1)
    for (int n2 = 0; n2 < n1; ++n2)
      at(n1, n2, n1) -= l2Factor * (c[n1] * c[n1] + c[n2] * c[n2]) / (nbrParticles - 1);
  const double lz = 0.5 * totalLz;
  shift = l2Factor * lz * (lz + 1.0);
}

ParticleOnSphereL2Hamiltonian::ParticleOnSphereL2Hamiltonian(FermionOnSphere* particles, int nbrParticles,
                                                             int lzmax, double l2Factor,
                                                             GPUArchitecture* architecture,
                                                             const std::vector<double>* pseudoPotential)
    : Particles(particles) {
  arch = architecture;
  if (particles->GetNbrParticles() != nbrParticles || particles->GetLzMax() != lzmax)
    throw std::invalid_argument("Hilbert space does not match the Hamiltonian parameters");
  double l2Shift = 0.0;
  EvaluateInteractionFactors(lzmax, nbrParticles, particles->GetTotalLz(), l2Factor, Factors, PairOffset, l2Shift);
  if (pseudoPotential) {
    if ((int)pseudoPotential->size() < lzmax + 1) throw std::invalid_argument("need lzMax + 1 pseudopotentials");
    std::vector<double> v;
    std::vector<int> off;
    ParticleOnSphereGenericHamiltonian::EvaluateInteractionFactors(lzmax, *pseudoPotential, v, off);
    for (size_t i = 0; i < v.size(); ++i) Factors[i] += v[i];
  }
  arch->Check(dh_sphere_hamiltonian_create(arch->Context(), &ham, particles->Basis(), Factors.data(),
                                           PairOffset.data(), (int)Factors.size()),
              "ParticleOnSphereL2Hamiltonian");
  ShiftHamiltonian(l2Shift);
}

// channels of the adjoint split: one per relative momentum m (odd), the table holds
// <m1 m2 | L> for every pair and L = lzMax - m
void ParticleOnSphereGenericHamiltonian::EvaluateChannelTable(int lzMax, std::vector<int>& channelM,
                                                              std::vector<double>& table) {
  const int nOrb = lzMax + 1;
  ClebschGordanCoefficients cg(lzMax, lzMax);
  channelM.clear();
  for (int m = 1; m <= lzMax; m += 2) channelM.push_back(m);
  const int nc = (int)channelM.size();
  const size_t nPairs = (size_t)nOrb * (nOrb - 1) / 2;
  table.assign(nPairs * nc, 0.0);
  for (int m1 = 0; m1 < nOrb; ++m1)
    for (int m2 = 0; m2 < m1; ++m2) {
      size_t a = (size_t)m1 * (m1 - 1) / 2 + m2;
      int d1 = 2 * m1 - lzMax, d2 = 2 * m2 - lzMax;
      for (int c = 0; c < nc; ++c) {
        int L = lzMax - channelM[c];
        table[a * nc + c] = cg.GetCoefficient(d1, d2, 2 * L);
      }
    }
}

ParticleOnSphereGenericHamiltonian::ParticleOnSphereGenericHamiltonian(FermionOnSphere* particles, int nbrParticles,
                                                                       int lzmax, const std::vector<double>& pseudo,
                                                                       GPUArchitecture* architecture)
    : Particles(particles), LzMax(lzmax), PseudoPotential(pseudo) {
  arch = architecture;
  if ((int)pseudo.size() < lzmax + 1) throw std::invalid_argument("need lzMax + 1 pseudopotentials");
  if (particles->GetNbrParticles() != nbrParticles || particles->GetLzMax() != lzmax)
    throw std::invalid_argument("Hilbert space does not match the Hamiltonian parameters");
  EvaluateInteractionFactors(lzmax, pseudo, Factors, PairOffset);
  arch->Check(dh_sphere_hamiltonian_create(arch->Context(), &ham, particles->Basis(), Factors.data(),
                                           PairOffset.data(), (int)Factors.size()),
              "ParticleOnSphereGenericHamiltonian");
  EvaluateChannelTable(lzmax, ChannelM, ChannelCG);
  std::vector<double> weights;
  for (int m : ChannelM) weights.push_back(pseudo[m]);
  arch->Check(dh_sphere_hamiltonian_set_channels(ham, ChannelCG.data(), (int)ChannelM.size(), weights.data()),
              "set channels");
}

ParticleOnSphereNBodyHamiltonian::ParticleOnSphereNBodyHamiltonian(FermionOnSphere* particles, int nbody,
                                                                   const std::vector<int>& sumCount,
                                                                   const std::vector<int>& tuples,
                                                                   const std::vector<double>& factors,
                                                                   GPUArchitecture* architecture)
    : Particles(particles), NBody(nbody) {
  arch = architecture;
  int64_t nt = 0, nf = 0;
  for (int c : sumCount) {
    nt += c;
    nf += (int64_t)c * c;
  }
  if ((int64_t)tuples.size() != nt * nbody || (int64_t)factors.size() != nf)
    throw std::invalid_argument("tuple or factor count does not match the per-sum tuple counts");
  arch->Check(dh_sphere_nbody_hamiltonian_create(arch->Context(), &ham, particles->Basis(), nbody, sumCount.data(),
                                                 (int)sumCount.size(), tuples.data(), factors.data()),
              "ParticleOnSphereNBodyHamiltonian");
}

void ParticleOnSphereGenericHamiltonian::Bilinears(const double* d_psi, const double* d_phi, double* g,
                                                   double* d_hpsi) {
  arch->Check(dh_sphere_adjoint(ham, d_psi, d_phi, g, d_hpsi), "Bilinears");
}

SpinChainHamiltonian::SpinChainHamiltonian(Spin1_2Chain* chain, int nbrSpin, const std::vector<double>& j,
                                           const std::vector<double>& jz, bool periodic,
                                           GPUArchitecture* architecture) {
  arch = architecture;
  int nb = periodic ? nbrSpin : nbrSpin - 1;
  if ((int)j.size() < nb || (int)jz.size() < nb) throw std::invalid_argument("not enough couplings");
  if (chain->GetChainLength() != nbrSpin) throw std::invalid_argument("chain length mismatch");
  arch->Check(dh_spin_chain_hamiltonian_create(arch->Context(), &ham, chain->Basis(), j.data(), jz.data(),
                                               periodic ? 1 : 0),
              "SpinChainHamiltonian");
}

}  // namespace dhh
