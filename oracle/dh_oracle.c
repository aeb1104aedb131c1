/*
 * dh_oracle.c -- CPU restatement of DiagHam's on-the-fly H.v + Lanczos path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load this library; the product path never
 * does.  Each function follows the reference file:line named in its comment
 * (paths relative to the DiagHam source tree).
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   basis order / dimension   FQHE/docs/DataFormat/FermionOnSphere.xml:23-37 (11-state
 *                             fixture), binomial / subset-sum closed forms, sortedness
 *   FindStateIndex            round trip find(state[i]) == i for every i
 *   signs / AdAdAA / AA+AdAd  brute-force Jordan-Wigner dense matrices (numpy, tests)
 *   Clebsch-Gordan            Racah closed form + orthonormality (tests)
 *   pseudopotentials          CG-sum restatement == closed form to 1e-13
 *   factors + H.v             reference's printed factor counts (3,432 / 1,661 at
 *                             2S=21) and full-matrix element count (1,016,988);
 *                             x.Hy == y.Hx; dense H vs second-quantised brute force
 *   Lanczos                   reference eigenvalues of configs N=8/2S=21 (10 values),
 *                             N=10/2S=27 (E0,E1), Heisenberg L=24 (4 values) to 1e-10;
 *                             numpy eigvalsh on small dense cases
 *   spin chain                Bethe-ansatz-free brute force (numpy dense) for L<=12,
 *                             reference L=24 eigenvalues
 */
#include "dh_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_MACHINE_PRECISION 1e-14 /* src/config.h:49 */

static inline uint64_t shr64(uint64_t x, int k) { return (k >= 64) ? 0ul : (x >> k); }

/* ======================================================================
 * Fermions on the sphere: FQHE/src/HilbertSpace/FermionOnSphere.cc
 * ====================================================================== */
typedef struct {
  int N, totalLz, lzMax, nbrLzValue;
  long dim;
  uint64_t* states;
  int* stateLzMax;
  int maxLookUpShift, lutMemorySize;
  int* lutShift;
  int** lut;
  double signTable[1 << 16];
  uint64_t signMask[128];
  uint64_t prodATmp;
  int prodALzMax;
} fs_space;

/* FermionOnSphere.cc:1809 ShiftedEvaluateHilbertSpaceDimension */
static long fs_shifted_dim(int n, int lzMax, int totalLz) {
  if ((n == 0) || (totalLz < 0) || (lzMax < (n - 1))) return 0l;
  int lzTotalMax = ((2 * lzMax - n + 1) * n) >> 1;
  if (lzTotalMax < totalLz) return 0l;
  if ((n == 1) && (lzMax >= totalLz)) return 1l;
  if (lzTotalMax == totalLz) return 1l;
  return fs_shifted_dim(n - 1, lzMax - 1, totalLz - lzMax) + fs_shifted_dim(n, lzMax - 1, totalLz);
}

/* FermionOnSphere.cc:1794 EvaluateHilbertSpaceDimension */
long or_fs_dimension(int nbrFermions, int lzMax, int totalLz) {
  return fs_shifted_dim(nbrFermions, lzMax, (totalLz + nbrFermions * lzMax) >> 1);
}

/* FermionOnSphere.cc:1621 GenerateStates (recursive, highest orbital first) */
static long fs_generate(fs_space* s, int n, int lzMax, int currentLzMax, int totalLz, long pos) {
  if ((n == 0) || (totalLz < 0) || (currentLzMax < (n - 1))) return pos;
  int lzTotalMax = ((2 * currentLzMax - n + 1) * n) >> 1;
  if (totalLz > lzTotalMax) return pos;
  if ((n == 1) && (currentLzMax >= totalLz)) {
    s->states[pos] = 0x1ul << totalLz;
    s->stateLzMax[pos] = lzMax;
    return pos + 1;
  }
  if (lzTotalMax == totalLz) {
    uint64_t mask = 0;
    for (int i = currentLzMax - n + 1; i <= currentLzMax; ++i) mask |= (0x1ul << i);
    s->states[pos] = mask;
    s->stateLzMax[pos] = lzMax;
    return pos + 1;
  }
  int reduced = currentLzMax - 1;
  long tmpPos = fs_generate(s, n - 1, lzMax, reduced, totalLz - currentLzMax, pos);
  uint64_t mask = 0x1ul << currentLzMax;
  for (long i = pos; i < tmpPos; i++) s->states[i] |= mask;
  if (lzMax == currentLzMax) return fs_generate(s, n, reduced, reduced, totalLz, tmpPos);
  return fs_generate(s, n, lzMax, reduced, totalLz, tmpPos);
}

/* FermionOnSphere.cc:1745 GenerateSignLookUpTable (64-bit masks) */
static void fs_sign_tables(fs_space* s) {
  for (int j = 0; j < (1 << 16); ++j) s->signTable[j] = (__builtin_popcount(j) & 1) ? -1.0 : 1.0;
  for (int i = 0; i < 48; ++i) s->signMask[i] = 0xfffful;
  for (int i = 48; i < 64; ++i) s->signMask[i] = 0xfffful >> (i - 48);
  for (int i = 64; i < 128; ++i) s->signMask[i] = 0x0ul;
}

/* FermionOnSphere.cc:1659 GenerateLookUpTable (one table per lzmax over the top bits) */
static void fs_lookup_table(fs_space* s, unsigned long memory) {
  memory /= (sizeof(int*) * s->nbrLzValue);
  s->maxLookUpShift = 1;
  while (memory > 0) {
    memory >>= 1;
    ++s->maxLookUpShift;
  }
  if (s->maxLookUpShift > s->nbrLzValue) s->maxLookUpShift = s->nbrLzValue;
  s->lutMemorySize = 1 << s->maxLookUpShift;
  s->lut = (int**)calloc(s->nbrLzValue, sizeof(int*));
  s->lutShift = (int*)calloc(s->nbrLzValue, sizeof(int));
  for (int i = 0; i < s->nbrLzValue; ++i) s->lut[i] = (int*)calloc(s->lutMemorySize + 1, sizeof(int));
  int currentLzMax = s->stateLzMax[0];
  int* tmpLut = s->lut[currentLzMax];
  s->lutShift[currentLzMax] = (currentLzMax < s->maxLookUpShift) ? 0 : currentLzMax + 1 - s->maxLookUpShift;
  int currentShift = s->lutShift[currentLzMax];
  unsigned long currentValue = s->lutMemorySize;
  unsigned long tmpValue = s->states[0] >> currentShift;
  while (currentValue > tmpValue) {
    tmpLut[currentValue] = 0;
    --currentValue;
  }
  tmpLut[currentValue] = 0;
  for (long i = 0; i < s->dim; ++i) {
    if (currentLzMax != s->stateLzMax[i]) {
      while (currentValue > 0) {
        tmpLut[currentValue] = (int)i;
        --currentValue;
      }
      tmpLut[0] = (int)i;
      currentLzMax = s->stateLzMax[i];
      tmpLut = s->lut[currentLzMax];
      s->lutShift[currentLzMax] = (currentLzMax < s->maxLookUpShift) ? 0 : currentLzMax + 1 - s->maxLookUpShift;
      currentShift = s->lutShift[currentLzMax];
      tmpValue = s->states[i] >> currentShift;
      currentValue = s->lutMemorySize;
      while (currentValue > tmpValue) {
        tmpLut[currentValue] = (int)i;
        --currentValue;
      }
      tmpLut[currentValue] = (int)i;
    } else {
      tmpValue = s->states[i] >> currentShift;
      if (tmpValue != currentValue) {
        while (currentValue > tmpValue) {
          tmpLut[currentValue] = (int)i;
          --currentValue;
        }
        tmpLut[currentValue] = (int)i;
      }
    }
  }
  while (currentValue > 0) {
    tmpLut[currentValue] = (int)(s->dim - 1);
    --currentValue;
  }
  tmpLut[0] = (int)(s->dim - 1);
  fs_sign_tables(s);
}

/* FermionOnSphere.cc:76 constructor (StateLzMax fix-up :108-115) */
void* or_fs_create(int nbrFermions, int totalLz, int lzMax, unsigned long memory) {
  fs_space* s = (fs_space*)calloc(1, sizeof(fs_space));
  s->N = nbrFermions;
  s->totalLz = totalLz;
  s->lzMax = lzMax;
  s->nbrLzValue = lzMax + 1;
  s->dim = (nbrFermions > 0) ? or_fs_dimension(nbrFermions, lzMax, totalLz) : 1;
  if (s->dim <= 0) {
    free(s);
    return NULL;
  }
  s->states = (uint64_t*)malloc(sizeof(uint64_t) * s->dim);
  s->stateLzMax = (int*)malloc(sizeof(int) * s->dim);
  if (nbrFermions > 0) {
    fs_generate(s, nbrFermions, lzMax, lzMax, (totalLz + nbrFermions * lzMax) >> 1, 0);
    if ((s->states[0] >> lzMax) == 0x0ul) {
      int tmp = lzMax;
      for (long i = 0; i < s->dim; ++i) {
        while ((s->states[i] >> tmp) == 0x0ul) --tmp;
        s->stateLzMax[i] = tmp;
      }
    }
  } else {
    s->states[0] = 0;
    s->stateLzMax[0] = 0;
  }
  fs_lookup_table(s, memory);
  return s;
}

void or_fs_destroy(void* p) {
  fs_space* s = (fs_space*)p;
  if (!s) return;
  for (int i = 0; i < s->nbrLzValue; ++i) free(s->lut[i]);
  free(s->lut);
  free(s->lutShift);
  free(s->states);
  free(s->stateLzMax);
  free(s);
}
long or_fs_dim(void* p) { return ((fs_space*)p)->dim; }
const uint64_t* or_fs_states(void* p) { return ((fs_space*)p)->states; }
const int* or_fs_state_lzmax(void* p) { return ((fs_space*)p)->stateLzMax; }

/* FermionOnSphere.cc:1431 FindStateIndex (LUT range + binary search; PosMin/PosMax as named there) */
long or_fs_find_state_index(void* p, uint64_t stateDescription, int lzmax) {
  fs_space* s = (fs_space*)p;
  long posMax = stateDescription >> s->lutShift[lzmax];
  long posMin = s->lut[lzmax][posMax];
  posMax = s->lut[lzmax][posMax + 1];
  long posMid = (posMin + posMax) >> 1;
  uint64_t current = s->states[posMid];
  while ((posMax != posMid) && (current != stateDescription)) {
    if (current > stateDescription)
      posMax = posMid;
    else
      posMin = posMid;
    posMid = (posMin + posMax) >> 1;
    current = s->states[posMid];
  }
  if (current == stateDescription) return posMid;
  return posMin;
}

/* sign factor of FermionOnSphere.cc:480-500: parity of the bits of (state >> n), 16 at a time */
static inline double fs_sign(const fs_space* s, uint64_t st, int n) {
  double c = s->signTable[shr64(st, n) & s->signMask[n]];
  c *= s->signTable[shr64(st, n + 16) & s->signMask[n + 16]];
  c *= s->signTable[shr64(st, n + 32) & s->signMask[n + 32]];
  c *= s->signTable[shr64(st, n + 48) & s->signMask[n + 48]];
  return c;
}

/* FermionOnSphere.cc:468 AdAdAA */
long or_fs_adadaa(void* p, long index, int m1, int m2, int n1, int n2, double* coefficient) {
  fs_space* s = (fs_space*)p;
  int stateLzMax = s->stateLzMax[index];
  uint64_t state = s->states[index];
  if ((n1 > stateLzMax) || (n2 > stateLzMax) || ((state & (0x1ul << n1)) == 0) || ((state & (0x1ul << n2)) == 0) ||
      (n1 == n2) || (m1 == m2)) {
    *coefficient = 0.0;
    return s->dim;
  }
  int newLzMax = stateLzMax;
  uint64_t tmp = state;
  double c = fs_sign(s, tmp, n2);
  tmp &= ~(0x1ul << n2);
  if (newLzMax == n2)
    while ((tmp >> newLzMax) == 0ul) --newLzMax;
  c *= fs_sign(s, tmp, n1);
  tmp &= ~(0x1ul << n1);
  if (tmp == 0x0ul)
    newLzMax = 0;
  else if (newLzMax == n1)
    while ((tmp >> newLzMax) == 0x0ul) --newLzMax;
  if ((tmp & (0x1ul << m2)) != 0x0ul) {
    *coefficient = 0.0;
    return s->dim;
  }
  if (m2 > newLzMax)
    newLzMax = m2;
  else
    c *= fs_sign(s, tmp, m2);
  tmp |= (0x1ul << m2);
  if ((tmp & (0x1ul << m1)) != 0x0ul) {
    *coefficient = 0.0;
    return s->dim;
  }
  if (m1 > newLzMax)
    newLzMax = m1;
  else
    c *= fs_sign(s, tmp, m1);
  tmp |= (0x1ul << m1);
  *coefficient = c;
  return or_fs_find_state_index(p, tmp, newLzMax);
}

/* FermionOnSphere.cc:790 AA: annihilate n2 then n1; the intermediate state and its highest
 * occupied orbital are what the reference caches in ProdATemporaryState / ProdALzMax */
static double fs_aa_local(const fs_space* s, long index, int n1, int n2, uint64_t* tmpOut, int* lzOut) {
  uint64_t tmp = s->states[index];
  if (((tmp & (0x1ul << n1)) == 0x0ul) || ((tmp & (0x1ul << n2)) == 0x0ul) || (n1 == n2)) {
    *tmpOut = tmp;
    return 0.0;
  }
  int lz = s->stateLzMax[index];
  double c = fs_sign(s, tmp, n2);
  tmp &= ~(0x1ul << n2);
  c *= fs_sign(s, tmp, n1);
  tmp &= ~(0x1ul << n1);
  if (tmp == 0x0ul) {
    lz = 0;
  } else {
    while ((tmp >> lz) == 0) --lz;
  }
  *tmpOut = tmp;
  *lzOut = lz;
  return c;
}

/* FermionOnSphere.cc:949 AdAd on the state produced by AA: create m2 then m1 */
static long fs_adad_local(const fs_space* s, uint64_t tmp, int lz, int m1, int m2, double* coefficient) {
  if ((tmp & (0x1ul << m2)) != 0x0ul) {
    *coefficient = 0.0;
    return s->dim;
  }
  int newLzMax = lz;
  double c = 1.0;
  if (m2 > newLzMax)
    newLzMax = m2;
  else
    c *= fs_sign(s, tmp, m2);
  tmp |= (0x1ul << m2);
  if ((tmp & (0x1ul << m1)) != 0x0ul) {
    *coefficient = 0.0;
    return s->dim;
  }
  if (m1 > newLzMax)
    newLzMax = m1;
  else
    c *= fs_sign(s, tmp, m1);
  tmp |= (0x1ul << m1);
  *coefficient = c;
  return or_fs_find_state_index((void*)s, tmp, newLzMax);
}

double or_fs_aa(void* p, long index, int n1, int n2) {
  fs_space* s = (fs_space*)p;
  return fs_aa_local(s, index, n1, n2, &s->prodATmp, &s->prodALzMax);
}

long or_fs_adad(void* p, int m1, int m2, double* coefficient) {
  fs_space* s = (fs_space*)p;
  return fs_adad_local(s, s->prodATmp, s->prodALzMax, m1, m2, coefficient);
}

/* FermionOnSphere.cc:716 ProdA: prod_i a_{n[i]} with the rightmost operator n[nbr-1] applied
 * first; the sign of each annihilation counts the occupied orbitals at and above it.  Returns
 * 0 if an orbital is empty; the intermediate state and its highest orbital go to *tmpOut, *lzOut. */
static double fs_prod_a_local(const fs_space* s, long index, const int* n, int nbr, uint64_t* tmpOut, int* lzOut) {
  uint64_t tmp = s->states[index];
  int lz = s->stateLzMax[index];
  double c = 1.0;
  for (int i = nbr - 1; i >= 0; --i) {
    int k = n[i];
    if ((tmp & (0x1ul << k)) == 0x0ul) {
      *tmpOut = tmp;
      *lzOut = lz;
      return 0.0;
    }
    c *= fs_sign(s, tmp, k);
    tmp &= ~(0x1ul << k);
  }
  if (tmp == 0x0ul)
    lz = 0;
  else
    while (((tmp >> lz) == 0) && (lz > 0)) --lz;
  *tmpOut = tmp;
  *lzOut = lz;
  return c;
}

/* FermionOnSphere.cc:875 ProdAd on the state kept by ProdA: prod_i a+_{m[i]}, rightmost first;
 * an orbital above the current highest one gives no sign.  Returns the target index (dim if an
 * orbital is already occupied). */
static long fs_prod_ad_local(const fs_space* s, uint64_t tmp, int lz, const int* m, int nbr, double* coefficient) {
  double c = 1.0;
  int newLz = lz;
  for (int i = nbr - 1; i >= 0; --i) {
    int k = m[i];
    if ((tmp & (0x1ul << k)) != 0x0ul) {
      *coefficient = 0.0;
      return s->dim;
    }
    if (k > newLz)
      newLz = k;
    else
      c *= fs_sign(s, tmp, k);
    tmp |= (0x1ul << k);
  }
  *coefficient = c;
  return or_fs_find_state_index((void*)s, tmp, newLz);
}

double or_fs_prod_a(void* p, long index, const int* n, int nbr) {
  fs_space* s = (fs_space*)p;
  return fs_prod_a_local(s, index, n, nbr, &s->prodATmp, &s->prodALzMax);
}

long or_fs_prod_ad(void* p, const int* m, int nbr, double* coefficient) {
  fs_space* s = (fs_space*)p;
  return fs_prod_ad_local(s, s->prodATmp, s->prodALzMax, m, nbr, coefficient);
}

/* AbstractQHEOnSphereNBodyInteractionHamiltonian::EvaluateMNNBodyAddMultiplyComponent with
 * explicit MN factors (AbstractQHEOnSphereNBodyInteractionHamiltonian.h:259-283), scatter form:
 * y[ProdAd(M_l)] += coef(ProdAd) * F_S[i][l] * ProdA(N_i) * x[index] for every index, every
 * orbital sum S, every annihilated tuple N_i and created tuple M_l of that sum.  Tuples are
 * nbody orbitals each, grouped by sum (sumCount[S] tuples of sum S, S = 0..nbrSums-1); the
 * factor matrices F_S (sumCount[S]^2, row i = annihilated tuple) follow each other. */
void or_nbody_hv(void* p, int nbody, int nbrSums, const int* sumCount, const int* tuples, const double* factors,
                 const double* x, double* y) {
  fs_space* s = (fs_space*)p;
  for (long index = 0; index < s->dim; ++index) {
    const int* T = tuples;
    const double* F = factors;
    for (int S = 0; S < nbrSums; ++S) {
      const int cnt = sumCount[S];
      for (int i = 0; i < cnt; ++i) {
        uint64_t tmp;
        int lz;
        double c3 = fs_prod_a_local(s, index, T + (size_t)i * nbody, nbody, &tmp, &lz);
        if (c3 == 0.0) continue;
        double c2 = c3 * x[index];
        for (int l = 0; l < cnt; ++l) {
          double c;
          long t = fs_prod_ad_local(s, tmp, lz, T + (size_t)l * nbody, nbody, &c);
          if (t < s->dim) y[t] += c * F[(size_t)i * cnt + l] * c2;
        }
      }
      T += (size_t)cnt * nbody;
      F += (size_t)cnt * cnt;
    }
  }
}

/* ======================================================================
 * Clebsch-Gordan coefficients: src/MathTools/ClebschGordanCoefficients.cc
 * (Schulten-Gordon recursion, :225 EvaluateClebschGordanCoefficients, :346 recursion)
 * ====================================================================== */
typedef struct {
  int J1, J2, totalM1, totalM2;
  int* jmin;     /* [totalM1*totalM2] */
  long* offset;  /* [totalM1*totalM2] start into coef */
  double* coef;
} cg_table;

#define CG_AT(t, p1, p2, j) ((t)->coef[(t)->offset[(p1) * (t)->totalM2 + (p2)] + (((j) - (t)->jmin[(p1) * (t)->totalM2 + (p2)]) >> 1)])

static double cg_C(const cg_table* t, int m1, int m2, int j) { /* :286 CCoefficient */
  double C = (double)(t->J2 - m2 + 2);
  C *= (double)(t->J2 + m2);
  C *= (double)(j + (m1 + m2));
  C *= (double)(j - (m1 + m2) + 2);
  return 0.25 * sqrt(C);
}
static double cg_D(const cg_table* t, int m1, int m2, int j) { /* :303 DCoefficient */
  return 0.25 * (double)(t->J2 * (t->J2 + 2) + j * (j + 2) - t->J1 * (t->J1 + 2) - 2 * (m2 * (m1 + m2)));
}

/* :315 RecursionAlgorithm */
static void cg_recursion(cg_table* t, int m1, int m2Min, int m2Max, int j) {
  int M2Mid = (m2Max + m2Min) / 2;
  int M1Pos = (m1 + t->J1) >> 1;
  int M2Pos = (m2Min + t->J2) >> 1;
  double prev = 1.0;
  double cur = -(cg_D(t, m1, m2Min, j) / cg_C(t, m1, m2Min + 2, j));
  double next = 0.0;
  int sgn = (t->J2 - m1 - j) / 2;
  if ((m2Max - m2Min) == 2) {
    double nf = 1.0 / sqrt((t->J1 + 1) * (1.0 + cur * cur));
    if ((((sgn & 1) == 0) && (cur < 0)) || (((sgn & 1) == 1) && (cur > 0))) nf *= -1.0;
    CG_AT(t, M1Pos, M2Pos, j) = nf;
    ++M2Pos;
    CG_AT(t, M1Pos, M2Pos, j) = cur * nf;
    return;
  }
  if ((m2Max - m2Min) == 4) {
    next = -((cg_D(t, m1, M2Mid, j) * cur + cg_C(t, m1, M2Mid, j) * prev) / cg_C(t, m1, m2Max, j));
    double nf = 1.0 / sqrt((t->J1 + 1) * (1.0 + next * next + cur * cur));
    if ((((sgn & 1) == 0) && (next < 0)) || (((sgn & 1) == 1) && (next > 0))) nf *= -1.0;
    CG_AT(t, M1Pos, M2Pos, j) = nf;
    ++M2Pos;
    CG_AT(t, M1Pos, M2Pos, j) = cur * nf;
    ++M2Pos;
    CG_AT(t, M1Pos, M2Pos, j) = next * nf;
    return;
  }
  if ((m2Max - m2Min) == 6) {
    next = -((cg_D(t, m1, m2Min + 2, j) * cur + cg_C(t, m1, m2Min + 2, j) * prev) / cg_C(t, m1, m2Max - 2, j));
    double last = -((cg_D(t, m1, m2Max - 2, j) * next + cg_C(t, m1, m2Max - 2, j) * cur) / cg_C(t, m1, m2Max, j));
    double nf = 1.0 / sqrt((t->J1 + 1) * (1.0 + next * next + cur * cur + last * last));
    if ((((sgn & 1) == 0) && (last < 0)) || (((sgn & 1) == 1) && (last > 0))) nf *= -1.0;
    CG_AT(t, M1Pos, M2Pos, j) = nf;
    ++M2Pos;
    CG_AT(t, M1Pos, M2Pos, j) = cur * nf;
    ++M2Pos;
    CG_AT(t, M1Pos, M2Pos, j) = next * nf;
    ++M2Pos;
    CG_AT(t, M1Pos, M2Pos, j) = last * nf;
    return;
  }
  if ((m2Max & 0x1) != (M2Mid & 0x1)) ++M2Mid;
  double norm = prev * prev;
  CG_AT(t, M1Pos, M2Pos, j) = prev;
  ++M2Pos;
  norm += cur * cur;
  CG_AT(t, M1Pos, M2Pos, j) = cur;
  ++M2Pos;
  for (int m2 = m2Min + 2; m2 < M2Mid; m2 += 2) {
    next = -((cg_D(t, m1, m2, j) * cur + cg_C(t, m1, m2, j) * prev) / cg_C(t, m1, m2 + 2, j));
    norm += next * next;
    CG_AT(t, M1Pos, M2Pos, j) = next;
    ++M2Pos;
    prev = cur;
    cur = next;
  }
  double mid = cur;
  if (fabs(mid) > (OR_MACHINE_PRECISION * norm)) {
    M2Pos = (m2Max + t->J2) >> 1;
    prev = 1.0;
    CG_AT(t, M1Pos, M2Pos, j) = prev;
    --M2Pos;
    cur = -(cg_D(t, m1, m2Max, j) / cg_C(t, m1, m2Max, j));
    CG_AT(t, M1Pos, M2Pos, j) = cur;
    --M2Pos;
    int reducedMid = M2Mid + 2;
    for (int m2 = m2Max - 2; m2 > reducedMid; m2 -= 2) {
      next = -((cg_D(t, m1, m2, j) * cur + cg_C(t, m1, m2 + 2, j) * prev) / cg_C(t, m1, m2, j));
      CG_AT(t, M1Pos, M2Pos, j) = next;
      --M2Pos;
      prev = cur;
      cur = next;
    }
    next = -((cg_D(t, m1, reducedMid, j) * cur + cg_C(t, m1, reducedMid + 2, j) * prev) / cg_C(t, m1, reducedMid, j));
    ++M2Pos;
    cur = mid / next;
    for (int m2 = reducedMid; m2 <= m2Max; m2 += 2) {
      CG_AT(t, M1Pos, M2Pos, j) *= cur;
      next = CG_AT(t, M1Pos, M2Pos, j);
      ++M2Pos;
      norm += next * next;
    }
  } else {
    for (int m2 = M2Mid; m2 < m2Max; m2 += 2) {
      next = -((cg_D(t, m1, m2, j) * cur + cg_C(t, m1, m2, j) * prev) / cg_C(t, m1, m2 + 2, j));
      norm += next * next;
      CG_AT(t, M1Pos, M2Pos, j) = next;
      ++M2Pos;
      prev = cur;
      cur = next;
    }
  }
  norm = 1.0 / sqrt(norm * (t->J1 + 1.0));
  if ((((sgn & 1) == 0) && (next < 0)) || (((sgn & 1) == 1) && (next > 0))) norm *= -1.0;
  M2Pos = (m2Min + t->J2) >> 1;
  for (int m2 = m2Min; m2 <= m2Max; m2 += 2) {
    CG_AT(t, M1Pos, M2Pos, j) *= norm;
    ++M2Pos;
  }
}

void* or_cg_create(int j1, int j2) {
  cg_table* t = (cg_table*)calloc(1, sizeof(cg_table));
  t->J1 = j1;
  t->J2 = j2;
  t->totalM1 = j1 + 1;
  t->totalM2 = j2 + 1;
  int sumJ = j1 + j2, diffJ = abs(j1 - j2);
  t->jmin = (int*)malloc(sizeof(int) * t->totalM1 * t->totalM2);
  t->offset = (long*)malloc(sizeof(long) * t->totalM1 * t->totalM2);
  long total = 0;
  for (int i = 0; i < t->totalM1; ++i)
    for (int j = 0; j < t->totalM2; ++j) {
      int sumM = abs(2 * (i + j) - sumJ);
      int jm = (sumM > diffJ) ? sumM : diffJ;
      t->jmin[i * t->totalM2 + j] = jm;
      t->offset[i * t->totalM2 + j] = total;
      total += sumJ - jm + 1;
    }
  t->coef = (double*)calloc(total, sizeof(double));
  for (int m1 = -j1; m1 <= j1; m1 += 2)
    for (int j = diffJ; j <= sumJ; j += 2) {
      int m2Max = j - m1;
      if (m2Max > j2) m2Max = j2;
      int m2Min = -j - m1;
      if (m2Min < -j2) m2Min = -j2;
      if (m2Max == m2Min) {
        int p1 = (m1 + j1) >> 1, p2 = (m2Min + j2) >> 1;
        int sgn = (j2 - m1 - j) / 2;
        CG_AT(t, p1, p2, j) = ((sgn & 1) == 0) ? 1.0 / sqrt(j1 + 1.0) : -1.0 / sqrt(j1 + 1.0);
      } else if (m2Max > m2Min) {
        cg_recursion(t, m1, m2Min, m2Max, j);
      }
    }
  double signCoef = ((j2 & 1) == 1) ? -1.0 : 1.0;
  for (int i = 0; i < t->totalM1; ++i) {
    double signCoef2 = signCoef;
    for (int j = 0; j < t->totalM2; ++j) {
      double* c = &t->coef[t->offset[i * t->totalM2 + j]];
      int jm = t->jmin[i * t->totalM2 + j];
      double J = (double)jm;
      int nb = (sumJ - jm) >> 1;
      for (int k = 0; k <= nb; ++k) {
        c[k] *= signCoef2 * sqrt(J + 1.0);
        J += 2.0;
      }
      signCoef2 *= -1.0;
    }
    signCoef *= -1.0;
  }
  return t;
}

void or_cg_destroy(void* p) {
  cg_table* t = (cg_table*)p;
  if (!t) return;
  free(t->jmin);
  free(t->offset);
  free(t->coef);
  free(t);
}

/* ClebschGordanCoefficients.cc:150 GetCoefficient (j must lie in [jmin, J1+J2]) */
double or_cg_get(void* p, int m1, int m2, int j) {
  cg_table* t = (cg_table*)p;
  if ((m1 >= -t->J1) && (m1 <= t->J1) && (m2 >= -t->J2) && (m2 <= t->J2)) {
    int p1 = (t->J1 + m1) >> 1, p2 = (t->J2 + m2) >> 1;
    int jm = t->jmin[p1 * t->totalM2 + p2];
    if (j < jm || j > t->J1 + t->J2 || ((j - jm) & 1)) return 0.0;
    return CG_AT(t, p1, p2, j);
  }
  return 0.0;
}

/* FQHE/src/Tools/FQHESpectrum/PseudoPotentials.cc:35 EvaluatePseudopotentials, LLL, d = 0 */
void or_coulomb_pseudopotentials(int nbrFlux, double* out) {
  int maxMomentum = nbrFlux;
  void* main = or_cg_create(maxMomentum, maxMomentum);
  void** coefs = (void**)malloc(sizeof(void*) * (maxMomentum + 1));
  for (int l = 0; l <= maxMomentum; ++l) coefs[l] = or_cg_create(maxMomentum, l << 1);
  /* layer separation 0: base = 1, every form factor is 1 */
  for (int l = 0; l <= maxMomentum; ++l) {
    double tmpPseudo = 0.0;
    for (int m1 = -maxMomentum; m1 <= maxMomentum; m1 += 2)
      for (int m2 = -maxMomentum; m2 <= maxMomentum; m2 += 2) {
        double tmpCoef = 0.0;
        int min = abs(m1 - m2) >> 1;
        double sign = 1.0;
        for (int j = min; j <= maxMomentum; ++j) {
          double c2 = or_cg_get(coefs[j], m1, m2 - m1, maxMomentum) * or_cg_get(coefs[j], nbrFlux, 0, maxMomentum);
          tmpCoef += sign * c2 * c2;
          sign *= -1.0;
        }
        tmpPseudo += or_cg_get(main, m1, -m1, l << 1) * or_cg_get(main, m2, -m2, l << 1) * tmpCoef;
      }
    out[maxMomentum - l] = tmpPseudo / sqrt(0.5 * nbrFlux);
  }
  for (int l = 0; l <= maxMomentum; ++l) or_cg_destroy(coefs[l]);
  free(coefs);
  or_cg_destroy(main);
}

/* FQHE/src/Hamiltonian/ParticleOnSphereCoulombHamiltonian.cc:213-223 (TmpV[j]) in the
 * pseudopotential-file order V_m, m = lzMax - j.  The file convention of
 * PseudoPotentials.cc:35 is 2 TmpV[j] / sqrt(S) (the Coulomb Hamiltonian folds the other
 * factor 2 into its own prefactor, :269); the tests pin this ratio against the CG sum. */
void or_coulomb_closed_form(int nbrFlux, double* out) {
  int lzMax = nbrFlux;
  for (int j = 0; j <= lzMax; ++j) {
    long double v = 1.0L;
    /* C(2L-2j, L-j) * C(2L+2j+2, L+j+1) / C(2L+2, L+1)^2 as interleaved ratios */
    int a = lzMax - j, b = lzMax + j + 1, c = lzMax + 1;
    long double num = 1.0L, den = 1.0L;
    for (int k = 1; k <= a; ++k) num *= (long double)(a + k) / (long double)k;
    for (int k = 1; k <= b; ++k) num *= (long double)(b + k) / (long double)k;
    for (int k = 1; k <= c; ++k) den *= (long double)(c + k) / (long double)k;
    v = num / (den * den);
    out[lzMax - j] = (double)(2.0L * v / sqrtl(0.5L * (long double)lzMax));
  }
}

/* ======================================================================
 * Interaction factors: FQHE/src/Hamiltonian/ParticleOnSphereGenericHamiltonian.cc:493
 * (fermionic branch) and AbstractQHEOnSphereHamiltonian.cc:165 (Hermitian symmetrisation)
 * ====================================================================== */
typedef struct {
  int lzMax, nbrLzValue, nbrM12;
  int *M1, *M2, *nbrM3, *M3off, *M3; /* M3 flattened, M3off prefix */
  int nbrFactors;
  double* F;
} or_factors_t;

void* or_factors_create(int lzMax, const double* pseudoPotentials) {
  or_factors_t* f = (or_factors_t*)calloc(1, sizeof(or_factors_t));
  int nbrLzValue = lzMax + 1;
  f->lzMax = lzMax;
  f->nbrLzValue = nbrLzValue;
  double* pp = (double*)malloc(sizeof(double) * nbrLzValue);
  for (int i = 0; i < nbrLzValue; ++i) pp[i] = pseudoPotentials[lzMax - i]; /* ctor :85 */
  void* cg = or_cg_create(lzMax, lzMax);
  double* tmp = (double*)calloc((size_t)nbrLzValue * nbrLzValue * nbrLzValue, sizeof(double));
  int sign = (lzMax & 1) ? 0 : 1;
  int pos = 0;
  for (int m1 = -lzMax; m1 <= lzMax; m1 += 2)
    for (int m2 = -lzMax; m2 < m1; m2 += 2) {
      int lim = m1 + m2 + lzMax;
      if (lim > lzMax) lim = lzMax;
      int min = m1 + m2 - lzMax;
      if (min < -lzMax) min = -lzMax;
      for (int m3 = min; m3 <= lim; m3 += 2) {
        int m4 = m1 + m2 - m3;
        tmp[pos] = 0.0;
        /* Clebsch iterator over J from |m1+m2| up to 2 lzMax in steps of 2 */
        int jmin = abs(m1 + m2);
        for (int J = jmin; J <= 2 * lzMax; J += 2)
          if (((J >> 1) & 1) == sign) tmp[pos] += pp[J >> 1] * or_cg_get(cg, m1, m2, J) * or_cg_get(cg, m3, m4, J);
        ++pos;
      }
    }
  f->F = (double*)malloc(sizeof(double) * (pos > 0 ? pos : 1));
  f->nbrFactors = 0;
  double factor = -4.0;
  f->nbrM12 = (nbrLzValue * (nbrLzValue - 1)) / 2;
  f->M1 = (int*)malloc(sizeof(int) * f->nbrM12);
  f->M2 = (int*)malloc(sizeof(int) * f->nbrM12);
  f->nbrM3 = (int*)malloc(sizeof(int) * f->nbrM12);
  f->M3off = (int*)malloc(sizeof(int) * (f->nbrM12 + 1));
  f->M3 = (int*)malloc(sizeof(int) * (pos > 0 ? pos : 1));
  int totalIndex = 0;
  pos = 0;
  int m3fill = 0;
  for (int m1 = 0; m1 < nbrLzValue; ++m1)
    for (int m2 = 0; m2 < m1; ++m2) {
      int lim = m1 + m2;
      if (lim > lzMax) lim = lzMax;
      int min = m1 + m2 - lzMax;
      if (min < 0) min = 0;
      f->M1[totalIndex] = m1;
      f->M2[totalIndex] = m2;
      f->nbrM3[totalIndex] = 0;
      f->M3off[totalIndex] = m3fill;
      for (int m3 = min; m3 <= lim; ++m3) {
        if ((2 * m3) > (m1 + m2)) {
          f->M3[m3fill++] = m3;
          f->F[f->nbrFactors++] = factor * tmp[pos];
          ++f->nbrM3[totalIndex];
        }
        ++pos;
      }
      ++totalIndex;
    }
  f->M3off[totalIndex] = m3fill;
  free(tmp);
  free(pp);
  or_cg_destroy(cg);
  return f;
}

/* FermionOnSphere.cc:1408 CheckOrder */
static int fs_check_order(const int* m, const int* n, int nbr) {
  uint64_t cv = 0, av = 0;
  for (int i = 0; i < nbr; ++i) {
    cv |= 0x1ul << m[i];
    av |= 0x1ul << n[i];
  }
  if (cv > av) return 1;
  if (cv < av) return -1;
  return 0;
}

/* AbstractQHEOnSphereHamiltonian.cc:165 HermitianSymmetrizeInteractionFactors (NbrM12Indices > 0) */
void or_factors_hermitian(void* p) {
  or_factors_t* f = (or_factors_t*)p;
  int* newM1 = (int*)malloc(sizeof(int) * f->nbrM12);
  int* newM2 = (int*)malloc(sizeof(int) * f->nbrM12);
  int* newNbr = (int*)malloc(sizeof(int) * f->nbrM12);
  int* newOff = (int*)malloc(sizeof(int) * (f->nbrM12 + 1));
  int* newM3 = (int*)malloc(sizeof(int) * (f->nbrFactors + 1));
  double* newF = (double*)malloc(sizeof(double) * (f->nbrFactors + 1));
  int nm12 = 0, nf = 0, pos = 0;
  for (int a = 0; a < f->nbrM12; ++a) {
    int M[2] = {f->M1[a], f->M2[a]};
    int sum = M[0] + M[1];
    int cnt = 0;
    int start = nf;
    for (int k = 0; k < f->nbrM3[a]; ++k) {
      int N[2] = {f->M3[f->M3off[a] + k], sum - f->M3[f->M3off[a] + k]};
      int flag = fs_check_order(M, N, 2);
      if (flag > 0) {
        newM3[nf] = N[0];
        newF[nf++] = f->F[pos];
        ++cnt;
      } else if (flag == 0) {
        newM3[nf] = N[0];
        newF[nf++] = 0.5 * f->F[pos];
        ++cnt;
      }
      ++pos;
    }
    if (cnt > 0) {
      newM1[nm12] = M[0];
      newM2[nm12] = M[1];
      newNbr[nm12] = cnt;
      newOff[nm12] = start;
      ++nm12;
    }
  }
  newOff[nm12] = nf;
  free(f->M1);
  free(f->M2);
  free(f->nbrM3);
  free(f->M3off);
  free(f->M3);
  free(f->F);
  f->M1 = newM1;
  f->M2 = newM2;
  f->nbrM3 = newNbr;
  f->M3off = newOff;
  f->M3 = newM3;
  f->F = newF;
  f->nbrM12 = nm12;
  f->nbrFactors = nf;
}

void or_factors_destroy(void* p) {
  or_factors_t* f = (or_factors_t*)p;
  if (!f) return;
  free(f->M1);
  free(f->M2);
  free(f->nbrM3);
  free(f->M3off);
  free(f->M3);
  free(f->F);
  free(f);
}
int or_factors_nbr_m12(void* p) { return ((or_factors_t*)p)->nbrM12; }
int or_factors_nbr(void* p) { return ((or_factors_t*)p)->nbrFactors; }
void or_factors_get(void* p, int* m1, int* m2, int* nbrM3, int* m3, double* F) {
  or_factors_t* f = (or_factors_t*)p;
  memcpy(m1, f->M1, sizeof(int) * f->nbrM12);
  memcpy(m2, f->M2, sizeof(int) * f->nbrM12);
  memcpy(nbrM3, f->nbrM3, sizeof(int) * f->nbrM12);
  memcpy(m3, f->M3, sizeof(int) * f->nbrFactors);
  memcpy(F, f->F, sizeof(double) * f->nbrFactors);
}

/* ======================================================================
 * H.v: AbstractQHEOnSphereHamiltonian.cc:410 LowLevelAddMultiply (NbrM12Indices > 0 branch)
 * and :2563 HermitianLowLevelAddMultiply
 * ====================================================================== */
static void fs_hv_range(const fs_space* sp, const or_factors_t* f, const double* x, double* y, double shift,
                        int hermitian, long first, long last) {
  long dim = sp->dim;
  for (long i = first; i < last; ++i) {
    int red = 0;
    double tmpSum = 0.0;
    for (int a = 0; a < f->nbrM12; ++a) {
      uint64_t tmp;
      int lz = 0;
      double c = fs_aa_local(sp, i, f->M1[a], f->M2[a], &tmp, &lz);
      if (c != 0.0) {
        int sum = f->M1[a] + f->M2[a];
        double c3 = c * x[i];
        const int* m3v = &f->M3[f->M3off[a]];
        for (int k = 0; k < f->nbrM3[a]; ++k) {
          double c2;
          long idx = fs_adad_local(sp, tmp, lz, m3v[k], sum - m3v[k], &c2);
          if (idx < dim) {
            if (hermitian) {
              tmpSum += x[idx] * c * f->F[red] * c2;
              y[idx] += f->F[red] * c2 * c3;
            } else {
              y[idx] += c3 * f->F[red] * c2;
            }
          }
          ++red;
        }
      } else {
        red += f->nbrM3[a];
      }
    }
    if (hermitian) y[i] += tmpSum;
  }
  for (long i = first; i < last; ++i) y[i] += shift * x[i];
}

/* the reference clones the basis per call because AA mutates it; the restatement keeps the
 * AA intermediate state in locals instead, so one basis serves every thread */
void or_fs_hv(void* space, void* factors, const double* x, double* y, double shift, int hermitian) {
  fs_space* s = (fs_space*)space;
  fs_hv_range(s, (or_factors_t*)factors, x, y, shift, hermitian, 0, s->dim);
}

/* VectorHamiltonianMultiplyOperation.cc:241-315: uniform row split, every thread but the
 * first writes into a private destination, the private destinations are summed at the end */
void or_fs_hv_smp(void* space, void* factors, const double* x, double* y, double shift, int hermitian,
                  int nbrThreads) {
  fs_space* s = (fs_space*)space;
  long dim = s->dim;
  if (nbrThreads < 1) nbrThreads = 1;
  double** priv = (double**)calloc(nbrThreads, sizeof(double*));
  for (int t = 1; t < nbrThreads; ++t) priv[t] = (double*)calloc(dim, sizeof(double));
#pragma omp parallel for num_threads(nbrThreads) schedule(static, 1)
  for (int t = 0; t < nbrThreads; ++t) {
    long step = dim / nbrThreads;
    long first = step * t;
    long last = (t == nbrThreads - 1) ? dim : first + step;
    fs_hv_range(s, (or_factors_t*)factors, x, (t == 0) ? y : priv[t], shift, hermitian, first, last);
  }
  for (int t = 1; t < nbrThreads; ++t) {
#pragma omp parallel for num_threads(nbrThreads)
    for (long i = 0; i < dim; ++i) y[i] += priv[t][i];
    free(priv[t]);
  }
  free(priv);
}

/* Non-Hermitian elements produced from source row `row`, in the reference loop order:
 * (target, AA sign * factor * AdAd sign).  These are also the elements of row `row` of H. */
long or_fs_row_elements(void* space, void* factors, long row, long* targets, double* coefs, long maxOut) {
  const fs_space* sp = (const fs_space*)space;
  or_factors_t* f = (or_factors_t*)factors;
  long n = 0;
  int red = 0;
  for (int a = 0; a < f->nbrM12; ++a) {
    uint64_t tmp;
    int lz = 0;
    double c = fs_aa_local(sp, row, f->M1[a], f->M2[a], &tmp, &lz);
    if (c != 0.0) {
      int sum = f->M1[a] + f->M2[a];
      const int* m3v = &f->M3[f->M3off[a]];
      for (int k = 0; k < f->nbrM3[a]; ++k) {
        double c2;
        long idx = fs_adad_local(sp, tmp, lz, m3v[k], sum - m3v[k], &c2);
        if (idx < sp->dim) {
          if (n < maxOut) {
            targets[n] = idx;
            coefs[n] = c * f->F[red] * c2;
          }
          ++n;
        }
        ++red;
      }
    } else {
      red += f->nbrM3[a];
    }
  }
  return n;
}

/* Gather form over rows [first, last) of a block of k vectors ([row][k] layout):
 * out[(r - first) k + v] = shift x[r k + v] + sum_t H[r][t] x[t k + v].  The elements of source
 * row r (or_fs_row_elements: targets t with H[t][r], in the reference loop order of
 * AbstractQHEOnSphereHamiltonian.cc:431-455) give row r of H because H is real symmetric.  A slice
 * of a large sector is checked this way without its whole product. */
void or_fs_hv_rows(void* space, void* factors, const double* x, double* out, double shift, long first, long last,
                   int k, int nbrThreads) {
  const or_factors_t* f = (const or_factors_t*)factors;
  const long cap = f->nbrFactors > 0 ? f->nbrFactors : 1;
  if (nbrThreads < 1) nbrThreads = 1;
#pragma omp parallel num_threads(nbrThreads)
  {
    long* t = (long*)malloc(sizeof(long) * cap);
    double* c = (double*)malloc(sizeof(double) * cap);
#pragma omp for schedule(dynamic, 64)
    for (long r = first; r < last; ++r) {
      long n = or_fs_row_elements(space, factors, r, t, c, cap);
      for (int v = 0; v < k; ++v) {
        double acc = shift * x[r * k + v];
        for (long e = 0; e < n; ++e) acc += c[e] * x[t[e] * k + v];
        out[(r - first) * k + v] = acc;
      }
    }
    free(t);
    free(c);
  }
}

/* The elements of every row stored once (CSR: row offsets, targets, coefficients), for solvers
 * that apply the same H many times; or_csr_hv is then the gather product of or_fs_hv_rows. */
typedef struct {
  long dim, nnz;
  long* rowOff;
  int* col;
  double* val;
} or_csr_t;

void* or_fs_csr_create(void* space, void* factors, int nbrThreads) {
  const fs_space* sp = (const fs_space*)space;
  const or_factors_t* f = (const or_factors_t*)factors;
  const long dim = sp->dim, cap = f->nbrFactors > 0 ? f->nbrFactors : 1;
  or_csr_t* m = (or_csr_t*)calloc(1, sizeof(or_csr_t));
  m->dim = dim;
  m->rowOff = (long*)calloc(dim + 1, sizeof(long));
  if (nbrThreads < 1) nbrThreads = 1;
#pragma omp parallel for num_threads(nbrThreads) schedule(dynamic, 256)
  for (long r = 0; r < dim; ++r) m->rowOff[r + 1] = or_fs_row_elements(space, factors, r, NULL, NULL, 0);
  for (long r = 0; r < dim; ++r) m->rowOff[r + 1] += m->rowOff[r];
  m->nnz = m->rowOff[dim];
  m->col = (int*)malloc(sizeof(int) * (m->nnz > 0 ? m->nnz : 1));
  m->val = (double*)malloc(sizeof(double) * (m->nnz > 0 ? m->nnz : 1));
#pragma omp parallel num_threads(nbrThreads)
  {
    long* t = (long*)malloc(sizeof(long) * cap);
    double* c = (double*)malloc(sizeof(double) * cap);
#pragma omp for schedule(dynamic, 256)
    for (long r = 0; r < dim; ++r) {
      long n = or_fs_row_elements(space, factors, r, t, c, cap);
      for (long e = 0; e < n; ++e) {
        m->col[m->rowOff[r] + e] = (int)t[e];
        m->val[m->rowOff[r] + e] = c[e];
      }
    }
    free(t);
    free(c);
  }
  return m;
}

void or_csr_destroy(void* p) {
  or_csr_t* m = (or_csr_t*)p;
  if (!m) return;
  free(m->rowOff);
  free(m->col);
  free(m->val);
  free(m);
}

long or_csr_nnz(void* p) { return ((or_csr_t*)p)->nnz; }

void or_csr_hv(void* p, const double* x, double* y, int nbrThreads) {
  const or_csr_t* m = (const or_csr_t*)p;
  if (nbrThreads < 1) nbrThreads = 1;
#pragma omp parallel for num_threads(nbrThreads) schedule(static, 4096)
  for (long r = 0; r < m->dim; ++r) {
    double acc = 0.0;
    for (long e = m->rowOff[r]; e < m->rowOff[r + 1]; ++e) acc += m->val[e] * x[m->col[e]];
    y[r] = acc;
  }
}

long or_fs_count_elements(void* space, void* factors) {
  fs_space* s = (fs_space*)space;
  long total = 0;
#pragma omp parallel for reduction(+ : total) schedule(dynamic, 256)
  for (long r = 0; r < s->dim; ++r) total += or_fs_row_elements(space, factors, r, NULL, NULL, 0);
  return total;
}

/* ======================================================================
 * Fermions in momentum space: FQHE/src/HilbertSpace/FermionOnSquareLatticeMomentumSpace.cc
 * (orbital index kx * Ny + ky, FermionOnSphere bit conventions, signs, LUT and search)
 * ====================================================================== */
typedef struct {
  int nx, ny, kx, ky;
} fm_params;

/* FermionOnSquareLatticeMomentumSpace.cc:349 EvaluateHilbertSpaceDimension */
static long fm_dimension(const fm_params* p, int n, int cKx, int cKy, int tKx, int tKy) {
  if (cKy < 0) {
    cKy = p->ny - 1;
    cKx--;
  }
  if (n == 0) return (((tKx % p->nx) == p->kx) && ((tKy % p->ny) == p->ky)) ? 1 : 0;
  if (cKx < 0) return 0;
  long count = 0;
  if (n == 1) {
    for (int j = cKy; j >= 0; --j)
      if ((((cKx + tKx) % p->nx) == p->kx) && (((j + tKy) % p->ny) == p->ky)) ++count;
    for (int i = cKx - 1; i >= 0; --i)
      for (int j = p->ny - 1; j >= 0; --j)
        if ((((i + tKx) % p->nx) == p->kx) && (((j + tKy) % p->ny) == p->ky)) ++count;
    return count;
  }
  count += fm_dimension(p, n - 1, cKx, cKy - 1, tKx + cKx, tKy + cKy);
  count += fm_dimension(p, n, cKx, cKy - 1, tKx, tKy);
  return count;
}

/* FermionOnSquareLatticeMomentumSpace.cc:290 GenerateStates */
static long fm_generate(const fm_params* p, uint64_t* st, int n, int cKx, int cKy, int tKx, int tKy, long pos) {
  if (cKy < 0) {
    cKy = p->ny - 1;
    cKx--;
  }
  if (n == 0) {
    if (((tKx % p->nx) == p->kx) && ((tKy % p->ny) == p->ky)) {
      st[pos] = 0x0ul;
      return pos + 1;
    }
    return pos;
  }
  if (cKx < 0) return pos;
  if (n == 1) {
    for (int j = cKy; j >= 0; --j)
      if ((((cKx + tKx) % p->nx) == p->kx) && (((j + tKy) % p->ny) == p->ky)) st[pos++] = 0x1ul << ((cKx * p->ny) + j);
    for (int i = cKx - 1; i >= 0; --i)
      for (int j = p->ny - 1; j >= 0; --j)
        if ((((i + tKx) % p->nx) == p->kx) && (((j + tKy) % p->ny) == p->ky)) st[pos++] = 0x1ul << ((i * p->ny) + j);
    return pos;
  }
  long tmp = fm_generate(p, st, n - 1, cKx, cKy - 1, tKx + cKx, tKy + cKy, pos);
  uint64_t mask = 0x1ul << ((cKx * p->ny) + cKy);
  for (; pos < tmp; ++pos) st[pos] |= mask;
  return fm_generate(p, st, n, cKx, cKy - 1, tKx, tKy, pos);
}

/* FermionOnSquareLatticeMomentumSpace.cc:76 constructor: LzMax = Nx * Ny */
void* or_fm_create(int nbrFermions, int nx, int ny, int kx, int ky, unsigned long memory) {
  fm_params p = {nx, ny, kx, ky};
  long dim = fm_dimension(&p, nbrFermions, nx - 1, ny - 1, 0, 0);
  if (dim <= 0) return NULL;
  fs_space* s = (fs_space*)calloc(1, sizeof(fs_space));
  s->N = nbrFermions;
  s->lzMax = nx * ny;
  s->nbrLzValue = s->lzMax + 1;
  s->dim = dim;
  s->states = (uint64_t*)malloc(sizeof(uint64_t) * dim);
  s->stateLzMax = (int*)malloc(sizeof(int) * dim);
  fm_generate(&p, s->states, nbrFermions, nx - 1, ny - 1, 0, 0, 0);
  int tmp = s->lzMax;
  for (long i = 0; i < dim; ++i) {
    while ((s->states[i] >> tmp) == 0x0ul) --tmp;
    s->stateLzMax[i] = tmp;
  }
  fs_lookup_table(s, memory);
  return s;
}

/* TightBindingModelCheckerboardLattice.cc:206 CoreComputeBandStructure (non-Bloch form):
 *   H(k) = [[d1 + d3, B1], [conj(B1), d1 - d3]],
 *   B1 = 4 t1 (cos(Kx/2) cos(Ky/2) cos(pi/4) + i sin(Kx/2) sin(Ky/2) sin(pi/4)),
 *   d1 = 4 t'' cos Kx cos Ky,  d3 = mu_s + 2 t2 (cos Kx - cos Ky),  K = 2 pi (k + gamma) / N.
 * The 2x2 problem is solved in closed form.  out_u[4 * index + (re0, im0, re1, im1)] holds the
 * lower-band one-body coefficients OneBodyBasis[index][0][alpha] = conj(v_alpha(k)), v the
 * normalised lower-band eigenvector with v_0 real >= 0 (v_1 real >= 0 when v_0 = 0); the
 * reference's own phase comes from HermitianMatrix::Diagonalize and is not restated (spectra
 * do not depend on it).  out_e[2 * index + band] = band energies. */
void or_checkerboard_lower_band(int nx, int ny, double t1, double t2, double tpp, double mus, double gx, double gy,
                                double* out_u, double* out_e) {
  for (int kx = 0; kx < nx; ++kx)
    for (int ky = 0; ky < ny; ++ky) {
      int index = kx * ny + ky;
      double Kx = 2.0 * M_PI * (kx + gx) / nx, Ky = 2.0 * M_PI * (ky + gy) / ny;
      double br = 4.0 * t1 * cos(0.5 * Kx) * cos(0.5 * Ky) * cos(M_PI * 0.25);
      double bi = 4.0 * t1 * sin(0.5 * Kx) * sin(0.5 * Ky) * sin(M_PI * 0.25);
      double d1 = 4.0 * tpp * cos(Kx) * cos(Ky);
      double d3 = mus + 2.0 * t2 * (cos(Kx) - cos(Ky));
      double r = sqrt(d3 * d3 + br * br + bi * bi);
      out_e[2 * index] = d1 - r;
      out_e[2 * index + 1] = d1 + r;
      double v0r, v0i, v1r, v1i;
      if (d3 >= 0.0) { /* (B1, -(d3 + r)) */
        v0r = br; v0i = bi; v1r = -(d3 + r); v1i = 0.0;
      } else { /* (r - d3, -conj(B1)) */
        v0r = r - d3; v0i = 0.0; v1r = -br; v1i = bi;
      }
      double nrm = sqrt(v0r * v0r + v0i * v0i + v1r * v1r + v1i * v1i);
      if (nrm == 0.0) { v0r = 1.0; nrm = 1.0; }
      v0r /= nrm; v0i /= nrm; v1r /= nrm; v1i /= nrm;
      double a0 = sqrt(v0r * v0r + v0i * v0i);
      double pr, pi; /* multiply by conj(phase) of the gauge-fixing component */
      if (a0 > 1e-300) { pr = v0r / a0; pi = -v0i / a0; }
      else { double a1 = sqrt(v1r * v1r + v1i * v1i); pr = v1r / a1; pi = -v1i / a1; }
      double w0r = v0r * pr - v0i * pi, w0i = v0r * pi + v0i * pr;
      double w1r = v1r * pr - v1i * pi, w1i = v1r * pi + v1i * pr;
      out_u[4 * index + 0] = w0r; out_u[4 * index + 1] = -w0i;
      out_u[4 * index + 2] = w1r; out_u[4 * index + 3] = -w1i;
    }
}

/* ParticleOnLatticeCheckerboardLatticeSingleBandHamiltonian.cc:126 EvaluateInteractionFactors
 * (fermions) with ComputeTwoBodyMatrixElementAB / AA (:359, :380) */
typedef struct {
  int nx, ny, nbrSums;
  int* nbrPerSum;  /* [nbrSums] */
  int** pairs;     /* [nbrSums][2 * nbrPerSum] (Index1 < Index2) */
  double** fre;    /* [nbrSums][nbrPerSum^2] */
  double** fim;
} fm_factors;

static void cmul(double ar, double ai, double br, double bi, double* cr, double* ci) {
  double r = ar * br - ai * bi, i = ar * bi + ai * br;
  *cr = r;
  *ci = i;
}

void* or_fm_factors_create(int nx, int ny, const double* u, double uPotential, double vPotential, int flatBand) {
  fm_factors* f = (fm_factors*)calloc(1, sizeof(fm_factors));
  f->nx = nx;
  f->ny = ny;
  f->nbrSums = nx * ny;
  f->nbrPerSum = (int*)calloc(f->nbrSums, sizeof(int));
  for (int kx1 = 0; kx1 < nx; ++kx1)
    for (int kx2 = 0; kx2 < nx; ++kx2)
      for (int ky1 = 0; ky1 < ny; ++ky1)
        for (int ky2 = 0; ky2 < ny; ++ky2) {
          int i1 = kx1 * ny + ky1, i2 = kx2 * ny + ky2;
          if (i1 < i2) ++f->nbrPerSum[(((kx1 + kx2) % nx) * ny) + ((ky1 + ky2) % ny)];
        }
  f->pairs = (int**)calloc(f->nbrSums, sizeof(int*));
  f->fre = (double**)calloc(f->nbrSums, sizeof(double*));
  f->fim = (double**)calloc(f->nbrSums, sizeof(double*));
  int* fill = (int*)calloc(f->nbrSums, sizeof(int));
  for (int i = 0; i < f->nbrSums; ++i) f->pairs[i] = (int*)malloc(sizeof(int) * 2 * (f->nbrPerSum[i] + 1));
  for (int kx1 = 0; kx1 < nx; ++kx1)
    for (int kx2 = 0; kx2 < nx; ++kx2)
      for (int ky1 = 0; ky1 < ny; ++ky1)
        for (int ky2 = 0; ky2 < ny; ++ky2) {
          int i1 = kx1 * ny + ky1, i2 = kx2 * ny + ky2;
          if (i1 < i2) {
            int sum = (((kx1 + kx2) % nx) * ny) + ((ky1 + ky2) % ny);
            f->pairs[sum][2 * fill[sum]] = i1;
            f->pairs[sum][2 * fill[sum] + 1] = i2;
            ++fill[sum];
          }
        }
  free(fill);
  double factorU = 0.5 / ((double)(nx * ny));
  if (!flatBand) factorU *= uPotential;
  double factorV = vPotential * 0.5 / ((double)(nx * ny));
  /* Kx = projected momenta (untilted lattice, no twist) */
#define KX(i) (2.0 * M_PI * ((i) / ny) / nx)
#define KY(i) (2.0 * M_PI * ((i) % ny) / ny)
#define UR(i, a) (u[4 * (i) + 2 * (a)])
#define UI(i, a) (u[4 * (i) + 2 * (a) + 1])
  for (int s = 0; s < f->nbrSums; ++s) {
    int n = f->nbrPerSum[s];
    f->fre[s] = (double*)calloc((size_t)n * n + 1, sizeof(double));
    f->fim[s] = (double*)calloc((size_t)n * n + 1, sizeof(double));
    int index = 0;
    for (int j1 = 0; j1 < n; ++j1) {
      int I1 = f->pairs[s][2 * j1], I2 = f->pairs[s][2 * j1 + 1];
      for (int j2 = 0; j2 < n; ++j2) {
        int I3 = f->pairs[s][2 * j2], I4 = f->pairs[s][2 * j2 + 1];
        /* four AB terms: (+,I1,I3,I2,I4 | k2,k4) (-,I2,I3,I1,I4 | k1,k4) (-,I1,I4,I2,I3 | k2,k3) (+,I2,I4,I1,I3 | k1,k3) */
        int ta[4][5] = {{+1, I1, I3, I2, I4}, {-1, I2, I3, I1, I4}, {-1, I1, I4, I2, I3}, {+1, I2, I4, I1, I3}};
        int tk[4][2] = {{I2, I4}, {I1, I4}, {I2, I3}, {I1, I3}};
        double accr = 0.0, acci = 0.0;
        for (int t = 0; t < 4; ++t) {
          int a = ta[t][1], b = ta[t][2], c = ta[t][3], d = ta[t][4];
          double pr, pi;
          cmul(UR(a, 0), -UI(a, 0), UR(b, 0), UI(b, 0), &pr, &pi);
          cmul(pr, pi, UR(c, 1), -UI(c, 1), &pr, &pi);
          cmul(pr, pi, UR(d, 1), UI(d, 1), &pr, &pi);
          int p1 = tk[t][0], p2 = tk[t][1];
          double m = 2.0 * (cos(0.5 * (KX(p2) - KY(p2) + KY(p1) - KX(p1))) + cos(0.5 * (KX(p2) + KY(p2) - KY(p1) - KX(p1))));
          accr += ta[t][0] * factorU * pr * m;
          acci += ta[t][0] * factorU * pi * m;
        }
        if (factorV != 0.0) {
          for (int t = 0; t < 4; ++t) {
            int a = ta[t][1], b = ta[t][2], c = ta[t][3], d = ta[t][4];
            double sr = 0.0, si = 0.0;
            for (int al = 0; al < 2; ++al) {
              double pr, pi;
              cmul(UR(a, al), -UI(a, al), UR(b, al), UI(b, al), &pr, &pi);
              cmul(pr, pi, UR(c, al), -UI(c, al), &pr, &pi);
              cmul(pr, pi, UR(d, al), UI(d, al), &pr, &pi);
              sr += pr;
              si += pi;
            }
            int p1 = tk[t][0], p2 = tk[t][1];
            double mr = cos(KX(p2) - KX(p1)) + cos(KY(p2) - KY(p1));
            double mi = sin(KX(p2) - KX(p1)) + sin(KY(p2) - KY(p1));
            double qr, qi;
            cmul(sr, si, mr, mi, &qr, &qi);
            accr += ta[t][0] * factorV * qr;
            acci += ta[t][0] * factorV * qi;
          }
        }
        f->fre[s][index] = -2.0 * accr;
        f->fim[s][index] = -2.0 * acci;
        ++index;
      }
    }
  }
#undef KX
#undef KY
#undef UR
#undef UI
  return f;
}

void or_fm_factors_destroy(void* p) {
  fm_factors* f = (fm_factors*)p;
  if (!f) return;
  for (int i = 0; i < f->nbrSums; ++i) {
    free(f->pairs[i]);
    free(f->fre[i]);
    free(f->fim[i]);
  }
  free(f->pairs);
  free(f->fre);
  free(f->fim);
  free(f->nbrPerSum);
  free(f);
}

int or_fm_factors_sector_size(void* p, int sum) { return ((fm_factors*)p)->nbrPerSum[sum]; }
void or_fm_factors_get(void* p, int sum, int* pairs, double* fre, double* fim) {
  fm_factors* f = (fm_factors*)p;
  int n = f->nbrPerSum[sum];
  memcpy(pairs, f->pairs[sum], sizeof(int) * 2 * n);
  memcpy(fre, f->fre[sum], sizeof(double) * n * n);
  memcpy(fim, f->fim[sum], sizeof(double) * n * n);
}

/* ParticleOnLatticeTimeReversalBreakingSingleBandHamiltonian.h:370
 * EvaluateMNTwoBodyAddMultiplyComponent (scatter) and :457 Hermitian variant; complex vectors are
 * interleaved (re, im).  Flat band: no one-body term. */
static void fm_hv_range(const fs_space* sp, const fm_factors* f, const double* x, double* y, int hermitian,
                        long first, long last) {
  long dim = sp->dim;
  for (long index = first; index < last; ++index) {
    double sumr = 0.0, sumi = 0.0;
    for (int j = 0; j < f->nbrSums; ++j) {
      int n = f->nbrPerSum[j];
      const int* idx = f->pairs[j];
      for (int i1 = 0; i1 < n; ++i1) {
        uint64_t tmp;
        int lz = 0;
        double c3 = fs_aa_local(sp, index, idx[2 * i1], idx[2 * i1 + 1], &tmp, &lz);
        if (c3 == 0.0) continue;
        double c4r = x[2 * index] * c3, c4i = x[2 * index + 1] * c3;
        const double* fr = &f->fre[j][i1 * n];
        const double* fi = &f->fim[j][i1 * n];
        for (int i2 = 0; i2 < n; ++i2) {
          double c;
          long target = fs_adad_local(sp, tmp, lz, idx[2 * i2], idx[2 * i2 + 1], &c);
          if (target >= dim) continue;
          if (hermitian) {
            if (target > index) continue;
            if (target < index) { /* TmpSum += x[target] * c * c3 * conj(F) */
              double ar = x[2 * target], ai = x[2 * target + 1];
              double pr, pi;
              cmul(ar, ai, fr[i2], -fi[i2], &pr, &pi);
              sumr += pr * c * c3;
              sumi += pi * c * c3;
            }
          }
          double pr, pi; /* y[target] += (c * F) * c4 */
          cmul(c * fr[i2], c * fi[i2], c4r, c4i, &pr, &pi);
          y[2 * target] += pr;
          y[2 * target + 1] += pi;
        }
      }
    }
    if (hermitian) {
      y[2 * index] += sumr;
      y[2 * index + 1] += sumi;
    }
  }
}

void or_fm_hv(void* space, void* factors, const double* x, double* y, int hermitian) {
  fs_space* s = (fs_space*)space;
  fm_hv_range(s, (fm_factors*)factors, x, y, hermitian, 0, s->dim);
}

/* elements (target, real and imaginary coefficient) produced from source row `row` */
long or_fm_row_elements(void* space, void* factors, long row, long* targets, double* cre, double* cim, long maxOut) {
  const fs_space* sp = (const fs_space*)space;
  const fm_factors* f = (const fm_factors*)factors;
  long cnt = 0;
  for (int j = 0; j < f->nbrSums; ++j) {
    int n = f->nbrPerSum[j];
    const int* idx = f->pairs[j];
    for (int i1 = 0; i1 < n; ++i1) {
      uint64_t tmp;
      int lz = 0;
      double c3 = fs_aa_local(sp, row, idx[2 * i1], idx[2 * i1 + 1], &tmp, &lz);
      if (c3 == 0.0) continue;
      for (int i2 = 0; i2 < n; ++i2) {
        double c;
        long target = fs_adad_local(sp, tmp, lz, idx[2 * i2], idx[2 * i2 + 1], &c);
        if (target >= sp->dim) continue;
        if (cnt < maxOut) {
          targets[cnt] = target;
          cre[cnt] = c * c3 * f->fre[j][i1 * n + i2];
          cim[cnt] = c * c3 * f->fim[j][i1 * n + i2];
        }
        ++cnt;
      }
    }
  }
  return cnt;
}

/* Gather form over rows [first, last) of a momentum sector for a block of k complex vectors
 * ([row][k][re, im]): out[((r - first) k + v)] = sum_t H[r][t] x[t][v] with H[r][t] = conj(H[t][r]),
 * the elements of source row r (or_fm_row_elements, the loop of
 * ParticleOnLatticeTimeReversalBreakingSingleBandHamiltonian.h:370) conjugated because H is
 * Hermitian. */
void or_fm_hv_rows(void* space, void* factors, const double* x, double* out, long first, long last, int k,
                   int nbrThreads) {
  const fm_factors* f = (const fm_factors*)factors;
  long cap = 1;
  for (int j = 0; j < f->nbrSums; ++j) cap += (long)f->nbrPerSum[j] * f->nbrPerSum[j];
  if (nbrThreads < 1) nbrThreads = 1;
#pragma omp parallel num_threads(nbrThreads)
  {
    long* t = (long*)malloc(sizeof(long) * cap);
    double* cr = (double*)malloc(sizeof(double) * cap);
    double* ci = (double*)malloc(sizeof(double) * cap);
#pragma omp for schedule(dynamic, 64)
    for (long r = first; r < last; ++r) {
      long n = or_fm_row_elements(space, factors, r, t, cr, ci, cap);
      for (int v = 0; v < k; ++v) {
        double ar = 0.0, ai = 0.0;
        for (long e = 0; e < n; ++e) {
          const double xr = x[2 * (t[e] * k + v)], xi = x[2 * (t[e] * k + v) + 1];
          ar += cr[e] * xr + ci[e] * xi;
          ai += cr[e] * xi - ci[e] * xr;
        }
        out[2 * ((r - first) * k + v)] = ar;
        out[2 * ((r - first) * k + v) + 1] = ai;
      }
    }
    free(t);
    free(cr);
    free(ci);
  }
}

/* ======================================================================
 * Fermions on the torus (one Ky sector, no magnetic translations):
 * FQHE/src/HilbertSpace/FermionOnTorus.cc (the momentum-space basis above with Nx = 1: orbital
 * ky, states of total momentum Ky mod Nphi in descending order, AA / AdAd identical to
 * FermionOnSphere's) and FQHE/src/Hamiltonian/ParticleOnTorusGenericHamiltonian.cc.
 * ====================================================================== */

/* L_n(x) by the three-term recurrence (the reference evaluates the monomial form of
 * LaguerrePolynomial(n)) */
static double torus_laguerre(int n, double x) {
  if (n == 0) return 1.0;
  double l0 = 1.0, l1 = 1.0 - x;
  for (int k = 1; k < n; ++k) {
    double l2 = ((2.0 * k + 1.0 - x) * l1 - k * l0) / (k + 1.0);
    l0 = l1;
    l1 = l2;
  }
  return l1;
}

typedef struct {
  int nflux, nbrPseudo, coulombLevel; /* coulombLevel < 0: pseudopotential form */
  double ratio;
  const double* pseudo;
} torus_params;

/* Fourier transform at q^2 * (2 pi / Nphi) = arg (pseudopotential form: sum_i V_i L_i(arg)
 * e^{-arg/2}), ParticleOnTorusGenericHamiltonian.cc:318 */
static double torus_vq_generic(const torus_params* p, double piOnM, double q2, int zeroQ) {
  double t = 0.0;
  for (int i = 0; i < p->nbrPseudo; ++i)
    if (p->pseudo[i] != 0.0) t += p->pseudo[i] * torus_laguerre(i, zeroQ ? 0.0 : 2.0 * piOnM * q2);
  return zeroQ ? t : exp(-piOnM * q2) * t;
}

/* ParticleOnTorusCoulombHamiltonian.cc:284 GetVofQ: L_n(q^2/2)^2 e^{-q^2/2} / |q| (0 at q = 0) */
static double torus_vq_coulomb(int level, double q2Half) {
  if (q2Half == 0.0) return 0.0;
  double ff = torus_laguerre(level, q2Half);
  return ff * ff / sqrt(2.0 * q2Half) * exp(-q2Half);
}

/* EvaluateInteractionCoefficient of ParticleOnTorusGenericHamiltonian.cc:318 (pseudopotentials) and
 * of ParticleOnTorusCoulombHamiltonian.cc:138: the lattice sum over q = (N1, N2) with
 * N2 = m1 - m4 mod Nphi, summed outwards in N2 (both directions) and N1 until the terms no longer
 * change the partial sums, the stopping rules of the reference kept as they are. */
static double torus_coefficient(const torus_params* p, int m1, int m2, int m3, int m4) {
  (void)m2;
  const double nbr = (double)p->nflux;
  const double piOnM = M_PI / nbr;
  const double factor = -((double)(m1 - m3)) * piOnM * 2.0;
  const int coul = p->coulombLevel >= 0;
  double sum = 0.0;
  for (int dir = 0; dir < 2; ++dir) {
    double n2 = dir == 0 ? (double)(m1 - m4) : (double)(m1 - m4 - p->nflux);
    double coefficient = dir == 0 ? 1.0 : sum;
    if (coul) coefficient = 1.0;
    while ((fabs(sum) + fabs(coefficient)) != fabs(sum)) {
      double n1 = 1.0;
      double q2 = p->ratio * n2 * n2;
      double precision;
      if (coul) {
        coefficient = torus_vq_coulomb(p->coulombLevel, piOnM * q2);
        precision = (n2 != 0.0) ? coefficient : 1.0;
      } else if (n2 != 0.0) {
        coefficient = torus_vq_generic(p, piOnM, q2, 0);
        precision = coefficient;
      } else {
        precision = 1.0;
        coefficient = torus_vq_generic(p, piOnM, 0.0, 1);
      }
      while ((fabs(coefficient) + precision) != fabs(coefficient)) {
        q2 = (1.0 / p->ratio) * n1 * n1 + p->ratio * n2 * n2;
        precision = 2.0 * (coul ? torus_vq_coulomb(p->coulombLevel, piOnM * q2) : torus_vq_generic(p, piOnM, q2, 0));
        coefficient += precision * cos(n1 * factor);
        n1 += 1.0;
      }
      sum += coefficient;
      n2 += dir == 0 ? nbr : -nbr;
    }
  }
  return coul ? sum / (2.0 * nbr) : sum / nbr;
}

/* ParticleOnTorusGenericHamiltonian::EvaluateInteractionFactors (:244, fermions) with the pair
 * lists of AbstractQHEOnTorusHamiltonian::GetIndices (:80): pairs m1 < m2 grouped by
 * (m1 + m2) mod Nphi; F[j1][j2] = V(m1,m2,m3,m4) + V(m2,m1,m4,m3) - V(m1,m2,m4,m3) - V(m2,m1,m3,m4)
 * with (m1, m2) = pair j1 (annihilated) and (m3, m4) = pair j2 (created).  Returned in the
 * momentum-space factor layout (Nx = 1, imaginary parts zero) for or_fm_hv. */
void* or_torus_factors_create(int nflux, double ratio, int nbrPseudo, const double* pseudo, int coulombLevel) {
  torus_params p = {nflux, nbrPseudo, coulombLevel, ratio, pseudo};
  fm_factors* f = (fm_factors*)calloc(1, sizeof(fm_factors));
  f->nx = 1;
  f->ny = nflux;
  f->nbrSums = nflux;
  f->nbrPerSum = (int*)calloc(nflux, sizeof(int));
  for (int m1 = 0; m1 < nflux - 1; ++m1)
    for (int m2 = m1 + 1; m2 < nflux; ++m2) ++f->nbrPerSum[(m1 + m2) % nflux];
  f->pairs = (int**)calloc(nflux, sizeof(int*));
  f->fre = (double**)calloc(nflux, sizeof(double*));
  f->fim = (double**)calloc(nflux, sizeof(double*));
  int* fill = (int*)calloc(nflux, sizeof(int));
  for (int s = 0; s < nflux; ++s) f->pairs[s] = (int*)malloc(sizeof(int) * 2 * (f->nbrPerSum[s] + 1));
  for (int m1 = 0; m1 < nflux - 1; ++m1)
    for (int m2 = m1 + 1; m2 < nflux; ++m2) {
      int s = (m1 + m2) % nflux;
      f->pairs[s][2 * fill[s]] = m1;
      f->pairs[s][2 * fill[s] + 1] = m2;
      ++fill[s];
    }
  free(fill);
  for (int s = 0; s < nflux; ++s) {
    int n = f->nbrPerSum[s];
    f->fre[s] = (double*)calloc((size_t)n * n + 1, sizeof(double));
    f->fim[s] = (double*)calloc((size_t)n * n + 1, sizeof(double));
    for (int j1 = 0; j1 < n; ++j1) {
      int m1 = f->pairs[s][2 * j1], m2 = f->pairs[s][2 * j1 + 1];
      for (int j2 = 0; j2 < n; ++j2) {
        int m3 = f->pairs[s][2 * j2], m4 = f->pairs[s][2 * j2 + 1];
        f->fre[s][j1 * n + j2] = torus_coefficient(&p, m1, m2, m3, m4) + torus_coefficient(&p, m2, m1, m4, m3) -
                                 torus_coefficient(&p, m1, m2, m4, m3) - torus_coefficient(&p, m2, m1, m3, m4);
      }
    }
  }
  return f;
}

/* ======================================================================
 * Fermions on the torus with magnetic translations:
 * FQHE/src/HilbertSpace/FermionOnTorusWithMagneticTranslations.cc (and the inline helpers of its
 * header) and the product of AbstractQHEOnTorusWithMagneticTranslationsHamiltonian.h:160-190.
 * ====================================================================== */
typedef struct {
  int N, nflux, kx, ky, modulo, shift, compShift;
  uint64_t mask;
  long dim;
  uint64_t* states;  /* canonical representatives, ascending */
  int* orbit;        /* NbrStateInOrbit */
  uint64_t* reorder; /* ReorderingSign: bit k = parity of the particles carried over k translations */
} ftt_space;

static int ftt_gcd(int a, int b) { return b == 0 ? a : ftt_gcd(b, a % b); }

/* translation by StateShift orbitals: the lowest `shift` orbitals wrap to the top */
static uint64_t ftt_translate(const ftt_space* s, uint64_t st) {
  return (st >> s->shift) | ((st & s->mask) << s->compShift);
}

/* FindCanonicalForm: the smallest state of the orbit, and the number of translations that take
 * the canonical state back to `st` */
static uint64_t ftt_canonical(const ftt_space* s, uint64_t st, int* nbrTranslation) {
  uint64_t best = st, cur = ftt_translate(s, st);
  int index = 1, at = 0;
  while (cur != st) {
    if (cur < best) {
      best = cur;
      at = index;
    }
    cur = ftt_translate(s, cur);
    ++index;
  }
  *nbrTranslation = at ? index - at : 0;
  return best;
}

/* FindNumberXTranslation and TestXMomentumConstraint */
static int ftt_orbit(const ftt_space* s, uint64_t st, int* signature) {
  uint64_t cur = st;
  int index = 0, sig = 0;
  do {
    sig += __builtin_popcountll(cur & s->mask) & 1;
    cur = ftt_translate(s, cur);
    ++index;
  } while (cur != st);
  if (signature) *signature = sig;
  return index;
}

static int ftt_kx_allowed(const ftt_space* s, uint64_t st) {
  int sig;
  int orbit = ftt_orbit(s, st, &sig);
  if (s->N & 1) return ((s->kx * orbit) % s->modulo) == 0;
  return (((s->kx * orbit) - ((s->modulo * sig) >> 1)) % s->modulo) == 0;
}

static int ftt_cmp(const void* a, const void* b) {
  uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* all raw states of total momentum ky (any order), then canonical forms, sorted ascending,
 * duplicates and states outside the Kx constraint removed (GenerateStates, :795) */
void* or_ftt_create(int nbrFermions, int nflux, int kx, int ky) {
  if (nbrFermions < 1 || nbrFermions > nflux || nflux > 62) return NULL;
  ftt_space* s = (ftt_space*)calloc(1, sizeof(ftt_space));
  s->N = nbrFermions;
  s->nflux = nflux;
  s->modulo = ftt_gcd(nbrFermions, nflux);
  s->kx = kx % s->modulo;
  s->ky = ky % nflux;
  s->shift = nflux / s->modulo;
  s->compShift = nflux - s->shift;
  s->mask = (1ul << s->shift) - 1ul;
  long cap = 1024, n = 0;
  uint64_t* raw = (uint64_t*)malloc(sizeof(uint64_t) * cap);
  /* subsets in colexicographic order by Gosper's hack */
  uint64_t st = (1ul << nbrFermions) - 1ul, limit = 1ul << nflux;
  while (st < limit) {
    int sum = 0;
    for (uint64_t t = st; t; t &= t - 1) sum += __builtin_ctzll(t);
    if (sum % nflux == s->ky) {
      int nt;
      if (n == cap) raw = (uint64_t*)realloc(raw, sizeof(uint64_t) * (cap *= 2));
      raw[n++] = ftt_canonical(s, st, &nt);
    }
    uint64_t c = st & (~st + 1), r = st + c;
    st = (((r ^ st) >> 2) / c) | r;
  }
  qsort(raw, n, sizeof(uint64_t), ftt_cmp);
  s->states = (uint64_t*)malloc(sizeof(uint64_t) * (n + 1));
  long d = 0;
  for (long i = 0; i < n; ++i) {
    if (i > 0 && raw[i] == raw[i - 1]) continue;
    if (!ftt_kx_allowed(s, raw[i])) continue;
    s->states[d++] = raw[i];
  }
  free(raw);
  s->dim = d;
  s->orbit = (int*)malloc(sizeof(int) * (d + 1));
  s->reorder = (uint64_t*)calloc(d + 1, sizeof(uint64_t));
  for (long i = 0; i < d; ++i) {
    s->orbit[i] = ftt_orbit(s, s->states[i], NULL);
    if ((s->N & 1) == 0) {
      uint64_t cur = s->states[i], bits = 0;
      for (int k = 1; k <= s->orbit[i]; ++k) {
        int moved = __builtin_popcountll(cur & s->mask) & 1;
        cur = ftt_translate(s, cur);
        uint64_t prev = (bits >> (k - 1)) & 1ul;
        bits |= (prev ^ (uint64_t)moved) << k;
      }
      s->reorder[i] = bits;
    }
  }
  return s;
}

void or_ftt_destroy(void* p) {
  ftt_space* s = (ftt_space*)p;
  if (!s) return;
  free(s->states);
  free(s->orbit);
  free(s->reorder);
  free(s);
}

long or_ftt_dim(void* p) { return ((ftt_space*)p)->dim; }
uint64_t* or_ftt_states(void* p) { return ((ftt_space*)p)->states; }
int* or_ftt_orbits(void* p) { return ((ftt_space*)p)->orbit; }
uint64_t* or_ftt_reorder(void* p) { return ((ftt_space*)p)->reorder; }
int or_ftt_modulo(void* p) { return ((ftt_space*)p)->modulo; }

static long ftt_find(const ftt_space* s, uint64_t st) {
  long lo = 0, hi = s->dim;
  while (lo < hi) {
    long mid = (lo + hi) >> 1;
    if (s->states[mid] < st)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < s->dim && s->states[lo] == st) ? lo : s->dim;
}

/* parity of the occupied orbitals at and above n, as the sign tables of the reference read it */
static double ftt_sign(uint64_t st, int n) { return (__builtin_popcountll(st >> n) & 1) ? -1.0 : 1.0; }

/* AdAdAA through AA (annihilate n2, then n1) and AdAd (create m2, then m1), :473 and :561 */
static long ftt_term(const ftt_space* s, long index, int n1, int n2, int m1, int m2, double* coefficient,
                     int* nbrTranslation) {
  uint64_t st = s->states[index];
  *coefficient = 0.0;
  if (!((st >> n1) & 1) || !((st >> n2) & 1) || n1 == n2 || m1 == m2) return s->dim;
  double c = ftt_sign(st, n2);
  st &= ~(1ul << n2);
  c *= ftt_sign(st, n1);
  st &= ~(1ul << n1);
  if ((st >> m2) & 1) return s->dim;
  c *= ftt_sign(st, m2); /* zero above m2 gives +1, as the reference's shortcut */
  st |= 1ul << m2;
  if ((st >> m1) & 1) return s->dim;
  c *= ftt_sign(st, m1);
  st |= 1ul << m1;
  int nt;
  uint64_t canon = ftt_canonical(s, st, &nt);
  if (!ftt_kx_allowed(s, canon)) return s->dim;
  long t = ftt_find(s, canon);
  if (t >= s->dim) return s->dim;
  c *= sqrt((double)s->orbit[index] / (double)s->orbit[t]);
  if ((s->reorder[t] >> nt) & 1ul) c = -c;
  *coefficient = c;
  *nbrTranslation = nt;
  return t;
}

/* y = H x over one (Kx, Ky) sector, complex interleaved, with the torus factor tables (or_torus_factors_create)
 * and the phases exp(2 pi i Kx t / modulo) of t translations (EvaluateExponentialFactors, :67) */
void or_ftt_hv(void* space, void* factors, const double* x, double* y) {
  const ftt_space* s = (const ftt_space*)space;
  const fm_factors* f = (const fm_factors*)factors;
  memset(y, 0, sizeof(double) * 2 * s->dim);
  for (long index = 0; index < s->dim; ++index) {
    const double xr = x[2 * index], xi = x[2 * index + 1];
    for (int j = 0; j < f->nbrSums; ++j) {
      int n = f->nbrPerSum[j];
      const int* idx = f->pairs[j];
      for (int i1 = 0; i1 < n; ++i1)
        for (int i2 = 0; i2 < n; ++i2) {
          double fr = f->fre[j][i1 * n + i2], fi = f->fim[j][i1 * n + i2];
          if (fr == 0.0 && fi == 0.0) continue;
          double c;
          int nt = 0;
          long t = ftt_term(s, index, idx[2 * i1], idx[2 * i1 + 1], idx[2 * i2], idx[2 * i2 + 1], &c, &nt);
          if (t >= s->dim) continue;
          double ph = 2.0 * M_PI * s->kx * (double)nt / (double)s->modulo;
          double er = cos(ph), ei = sin(ph);
          double ar, ai, br, bi;
          cmul(c * fr, c * fi, er, ei, &ar, &ai);
          cmul(ar, ai, xr, xi, &br, &bi);
          y[2 * t] += br;
          y[2 * t + 1] += bi;
        }
    }
  }
}

/* ======================================================================
 * Spin-1/2 chain: Spin/src/HilbertSpace/Spin1_2Chain.cc, Spin/src/Hamiltonian/SpinChainHamiltonian.cc
 * ====================================================================== */
typedef struct {
  int L, sz;
  long dim;
  uint64_t* states;
  int* lut;
  int lutSize, lookUpPosition;
  uint64_t lutMask;
} spin_space;

/* Spin1_2Chain.cc:331 GenerateStates (site 0 decided first; bits >= L stay set) */
static int spin_generate(spin_space* s, int statePosition, int sitePosition, unsigned currentState, int currentSz) {
  int maxSz = (s->L - sitePosition) * 2;
  int diffSz = currentSz - s->sz;
  if (diffSz > maxSz) return 0;
  int nbrGenerated = -statePosition;
  int next = sitePosition + 1;
  uint64_t mask;
  sitePosition &= 0x1f;
  if (next != s->L) {
    if (next == s->lookUpPosition) s->lut[currentState & s->lutMask] = statePosition;
    statePosition += spin_generate(s, statePosition, next, currentState, currentSz);
    mask = ~(0x1ul << sitePosition);
    if (next == s->lookUpPosition) s->lut[(currentState & mask) & s->lutMask] = statePosition;
    statePosition += spin_generate(s, statePosition, next, (unsigned)(currentState & mask), currentSz - 2);
  } else {
    if (currentSz == s->sz) {
      if (next == s->lookUpPosition) s->lut[currentState & s->lutMask] = statePosition;
      s->states[statePosition] = currentState;
      statePosition++;
    } else {
      currentSz -= 2;
      if (currentSz == s->sz) {
        mask = ~(0x1ul << sitePosition);
        if (next == s->lookUpPosition) s->lut[(currentState & mask) & s->lutMask] = statePosition;
        s->states[statePosition] = (currentState & mask);
        statePosition++;
      }
    }
  }
  nbrGenerated += statePosition;
  return nbrGenerated;
}

static long binom_l(int n, int k) {
  if (k < 0 || k > n) return 0;
  long r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

/* Spin1_2Chain.cc:113 constructor with fixed Sz */
void* or_spin_create(int chainLength, int sz, int memorySize) {
  spin_space* s = (spin_space*)calloc(1, sizeof(spin_space));
  s->L = chainLength;
  if (s->L > 32) s->L = 1;
  s->sz = sz;
  s->dim = binom_l(s->L, (s->L + sz) / 2); /* :1039 EvaluateHilbertSpaceDimension */
  s->lookUpPosition = 0;
  s->lutSize = 1;
  memorySize >>= 2;
  uint64_t mask = 0xfffffffful;
  while ((s->lookUpPosition < s->L) && (memorySize >= 2)) {
    mask <<= 1;
    s->lutSize *= 2;
    memorySize >>= 1;
    s->lookUpPosition++;
  }
  s->lutMask = (~mask) & 0xffffffffffffffffull;
  /* the reference keeps the mask in a 32-bit-seeded unsigned long: only the low bits matter */
  s->lutMask &= (s->lutSize - 1);
  s->lut = (int*)calloc(s->lutSize, sizeof(int));
  s->states = (uint64_t*)malloc(sizeof(uint64_t) * (s->dim > 0 ? s->dim : 1));
  s->states[0] = 0xfffffffful;
  s->dim = spin_generate(s, 0, 0, 0xffffffffu, s->L);
  return s;
}
void or_spin_destroy(void* p) {
  spin_space* s = (spin_space*)p;
  if (!s) return;
  free(s->states);
  free(s->lut);
  free(s);
}
long or_spin_dim(void* p) { return ((spin_space*)p)->dim; }
const uint64_t* or_spin_states(void* p) { return ((spin_space*)p)->states; }

/* Spin1_2Chain.cc:1002 FindStateIndex (LUT on the low bits, then linear scan) */
long or_spin_find_state_index(void* p, uint64_t state) {
  spin_space* s = (spin_space*)p;
  long index = s->lut[state & s->lutMask];
  const uint64_t* tmp = &s->states[index];
  while ((index < s->dim) && (state != *(tmp++))) index++;
  return index;
}

/* Spin1_2Chain.cc:655 SmiSpj for i != j */
static inline long spin_smispj(spin_space* s, int i, int j, long state, double* coef) {
  uint64_t st = s->states[state];
  uint64_t bi = (st >> i) & 0x1ul, bj = (st >> j) & 0x1ul;
  if (((bj << 1) | bi) == 0x1ul) {
    *coef = 1.0;
    return or_spin_find_state_index(s, (st | (0x1ul << j)) & ~(0x1ul << i));
  }
  *coef = 0.0;
  return s->dim;
}

/* Spin1_2Chain.cc:561 SziSzj */
static inline double spin_szisz(const spin_space* s, int i, int j, long state) {
  uint64_t mask = (0x1ul << i) | (0x1ul << j);
  uint64_t t = s->states[state] & mask;
  return ((t == 0) || (t == mask)) ? 0.25 : -0.25;
}

/* SpinChainHamiltonian.cc:257 LowLevelAddMultiply with :346 EvaluateDiagonalMatrixElements */
static void spin_hv_range(spin_space* s, const double* J, const double* Jz, int periodic, const double* x, double* y,
                          long first, long last) {
  int maxPos = s->L - 1;
  for (long i = first; i < last; ++i) {
    double szsz = 0.0;
    for (int j = 0; j < maxPos; ++j) szsz += Jz[j] * spin_szisz(s, j, j + 1, i);
    if (periodic) szsz += Jz[maxPos] * spin_szisz(s, maxPos, 0, i);
    double v = x[i];
    y[i] += szsz * v;
    double coef;
    long pos;
    for (int j = 0; j < maxPos; ++j) {
      pos = spin_smispj(s, j, j + 1, i, &coef);
      if (pos != s->dim) y[pos] += 0.5 * J[j] * coef * v;
      pos = spin_smispj(s, j + 1, j, i, &coef);
      if (pos != s->dim) y[pos] += 0.5 * J[j] * coef * v;
    }
    if (periodic) {
      pos = spin_smispj(s, maxPos, 0, i, &coef);
      if (pos != s->dim) y[pos] += 0.5 * J[maxPos] * coef * v;
      pos = spin_smispj(s, 0, maxPos, i, &coef);
      if (pos != s->dim) y[pos] += 0.5 * J[maxPos] * coef * v;
    }
  }
}

void or_spin_hv(void* space, const double* J, const double* Jz, int periodic, const double* x, double* y) {
  spin_space* s = (spin_space*)space;
  spin_hv_range(s, J, Jz, periodic, x, y, 0, s->dim);
}

/* rows [first, last) only, split over threads like or_spin_hv_smp: the bounded sample the
 * CPU baseline times */
void or_spin_hv_rows(void* space, const double* J, const double* Jz, int periodic, const double* x, double* y,
                     long first, long last, int nbrThreads) {
  spin_space* s = (spin_space*)space;
  long dim = s->dim;
  if (nbrThreads < 1) nbrThreads = 1;
  double** priv = (double**)calloc(nbrThreads, sizeof(double*));
  for (int t = 1; t < nbrThreads; ++t) priv[t] = (double*)calloc(dim, sizeof(double));
#pragma omp parallel for num_threads(nbrThreads) schedule(static, 1)
  for (int t = 0; t < nbrThreads; ++t) {
    long step = (last - first) / nbrThreads;
    long a = first + step * t;
    long b = (t == nbrThreads - 1) ? last : a + step;
    spin_hv_range(s, J, Jz, periodic, x, (t == 0) ? y : priv[t], a, b);
  }
  for (int t = 1; t < nbrThreads; ++t) {
#pragma omp parallel for num_threads(nbrThreads)
    for (long i = 0; i < dim; ++i) y[i] += priv[t][i];
    free(priv[t]);
  }
  free(priv);
}

void or_spin_hv_smp(void* space, const double* J, const double* Jz, int periodic, const double* x, double* y,
                    int nbrThreads) {
  spin_space* s = (spin_space*)space;
  long dim = s->dim;
  if (nbrThreads < 1) nbrThreads = 1;
  double** priv = (double**)calloc(nbrThreads, sizeof(double*));
  for (int t = 1; t < nbrThreads; ++t) priv[t] = (double*)calloc(dim, sizeof(double));
#pragma omp parallel for num_threads(nbrThreads) schedule(static, 1)
  for (int t = 0; t < nbrThreads; ++t) {
    long step = dim / nbrThreads;
    long first = step * t;
    long last = (t == nbrThreads - 1) ? dim : first + step;
    spin_hv_range(s, J, Jz, periodic, x, (t == 0) ? y : priv[t], first, last);
  }
  for (int t = 1; t < nbrThreads; ++t) {
#pragma omp parallel for num_threads(nbrThreads)
    for (long i = 0; i < dim; ++i) y[i] += priv[t][i];
    free(priv[t]);
  }
  free(priv);
}

/* ======================================================================
 * Lanczos: src/LanczosAlgorithm/FullReorthogonalizedLanczosAlgorithm.cc:122-360 and
 * src/Matrix/RealTriDiagonalSymmetricMatrix.cc:687 Diagonalize
 * ====================================================================== */

/* RealTriDiagonalSymmetricMatrix.cc:687 (QL, implicit shift); eigenvalues left in diag */
int or_tridiag_eigenvalues(int n, double* d, double* u, int maxIter) {
  int red = n - 1;
  for (int i = 0; i < red; i++) {
    int iter = 0;
    while (iter < maxIter) {
      int j = i;
      int flag = 0;
      while ((j < red) && (flag == 0)) {
        double d2 = fabs(d[j]) + fabs(d[j + 1]);
        if ((d2 + fabs(u[j])) == d2)
          flag = 1;
        else
          j++;
      }
      if (j != i) {
        iter++;
        double theta = (d[i + 1] - d[i]) / (2.0 * u[i]);
        double R = sqrt(1.0 + theta * theta);
        double T = d[j] - d[i];
        if (theta >= 0)
          T += u[i] / (theta + R);
        else
          T += u[i] / (theta - R);
        double Cos = 1.0, Sin = 1.0, P = 0.0, F, B;
        for (int k = j - 1; k >= i; k--) {
          F = Sin * u[k];
          B = Cos * u[k];
          R = sqrt(F * F + T * T);
          u[k + 1] = R;
          if (R == 0.0) {
            d[k + 1] -= P;
            u[j] = 0.0;
            k = i - 1;
          } else {
            Sin = 1.0 / R;
            Cos = Sin * T;
            Sin *= F;
            T = d[k + 1] - P;
            R = (d[k] - T) * Sin + 2.0 * Cos * B;
            P = Sin * R;
            d[k + 1] = T + P;
            T = Cos * R - B;
          }
        }
        d[i] -= P;
        u[i] = T;
        u[j] = 0.0;
      } else
        iter = maxIter;
    }
  }
  return 0;
}

static int cmp_double(const void* a, const void* b) {
  double x = *(const double*)a, y = *(const double*)b;
  return (x < y) ? -1 : (x > y);
}

static uint64_t splitmix_next(uint64_t* state) {
  uint64_t z = (*state += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

/* uniform in [-1, 1) from SplitMix64 (seeded inputs shared by tests and bench) */
void or_splitmix_fill(uint64_t seed, double* out, long n) {
  uint64_t st = seed;
  for (long i = 0; i < n; ++i) {
    uint64_t r = splitmix_next(&st);
    out[i] = ((double)(r >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
  }
}

typedef void (*or_hv_fn)(void* ctx, const double* x, double* y);

static double dotp(const double* a, const double* b, long n) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s)
  for (long i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* FullReorthogonalizedLanczosAlgorithm: Run(nbrEigen + 2) then Run(1) until converged
 * (QHEOnSphereMainTask.cc:732-753); strong convergence test :325 on every wanted value */
static int or_lanczos_core(or_hv_fn hv, void* ctx, long dim, int nbrEigen, int maxIter, double precision,
                           uint64_t seed, double* eigenvalues) {
  if (maxIter < nbrEigen + 3) maxIter = nbrEigen + 3;
  double** V = (double**)calloc(maxIter + 2, sizeof(double*));
  double* alpha = (double*)calloc(maxIter + 2, sizeof(double));
  double* beta = (double*)calloc(maxIter + 2, sizeof(double));
  double* prev = (double*)calloc(nbrEigen, sizeof(double));
  double* work_d = (double*)malloc(sizeof(double) * (maxIter + 2));
  double* work_u = (double*)malloc(sizeof(double) * (maxIter + 2));
  uint64_t rng = seed;
  for (int i = 0; i < 3; ++i) V[i] = (double*)calloc(dim, sizeof(double));
  or_splitmix_fill(splitmix_next(&rng), V[0], dim);
  double nrm = sqrt(dotp(V[0], V[0], dim));
  for (long k = 0; k < dim; ++k) V[0][k] /= nrm;
  /* first two steps (Index == 0 branch, :213-236) */
  hv(ctx, V[0], V[1]);
  alpha[0] = dotp(V[0], V[1], dim);
  for (long k = 0; k < dim; ++k) V[1][k] -= alpha[0] * V[0][k];
  nrm = sqrt(dotp(V[1], V[1], dim));
  for (long k = 0; k < dim; ++k) V[1][k] /= nrm;
  hv(ctx, V[1], V[2]);
  beta[0] = dotp(V[0], V[2], dim);
  alpha[1] = dotp(V[1], V[2], dim);
  int index = 0;
  int converged = 0;
  int nbrRows = 2;
  int target = nbrEigen + 2;
  int first = 1;
  while (!converged) {
    while (nbrRows < target) {
      int i = index + 2;
      if (i + 1 >= maxIter + 2) break;
      for (long k = 0; k < dim; ++k) V[i][k] -= alpha[index + 1] * V[i - 1][k] + beta[index] * V[i - 2][k];
      if (i > 2) { /* classical Gram-Schmidt against v_0..v_{i-3} (:256-269) */
        double* c = (double*)malloc(sizeof(double) * i);
        for (int j = 0; j < i - 2; ++j) c[j] = dotp(V[i], V[j], dim);
        for (int j = 0; j < i - 2; ++j) {
          double cj = c[j];
          const double* vj = V[j];
#pragma omp parallel for
          for (long k = 0; k < dim; ++k) V[i][k] -= cj * vj[k];
        }
        free(c);
      }
      double vn = sqrt(dotp(V[i], V[i], dim));
      while (vn < 1e-5) { /* subspace collapse (:270-296): random restart orthogonal to all */
        or_splitmix_fill(splitmix_next(&rng), V[i], dim);
        double* tmpv = (double*)calloc(dim, sizeof(double));
        hv(ctx, V[i], tmpv);
        memcpy(V[i], tmpv, sizeof(double) * dim);
        free(tmpv);
        for (int j = 0; j < i; ++j) {
          double cj = dotp(V[j], V[i], dim);
          for (long k = 0; k < dim; ++k) V[i][k] -= cj * V[j][k];
        }
        vn = sqrt(dotp(V[i], V[i], dim));
      }
      for (long k = 0; k < dim; ++k) V[i][k] /= vn;
      index++;
      V[i + 1] = (double*)calloc(dim, sizeof(double));
      hv(ctx, V[i], V[i + 1]);
      beta[index] = dotp(V[i - 1], V[i + 1], dim);
      alpha[index + 1] = dotp(V[i], V[i + 1], dim);
      nbrRows++;
    }
    memcpy(work_d, alpha, sizeof(double) * nbrRows);
    memcpy(work_u, beta, sizeof(double) * nbrRows);
    or_tridiag_eigenvalues(nbrRows, work_d, work_u, 50);
    qsort(work_d, nbrRows, sizeof(double), cmp_double);
    if (first) {
      for (int e = 0; e < nbrEigen; ++e) prev[e] = 2.0 * work_d[e];
      first = 0;
    } else {
      converged = (nbrRows > nbrEigen);
      for (int e = nbrEigen - 1; e >= 0 && converged; --e)
        if (fabs(work_d[e] - prev[e]) > precision * fabs(work_d[e])) {
          if (fabs(work_d[e]) > 3 * OR_MACHINE_PRECISION || fabs(prev[e]) > 3 * OR_MACHINE_PRECISION) converged = 0;
        }
      for (int e = 0; e < nbrEigen; ++e) prev[e] = work_d[e];
    }
    if (nbrRows + 1 >= maxIter) break;
    target = nbrRows + 1;
  }
  for (int e = 0; e < nbrEigen; ++e) eigenvalues[e] = work_d[e];
  int iterations = nbrRows;
  for (int i = 0; i < maxIter + 2; ++i) free(V[i]);
  free(V);
  free(alpha);
  free(beta);
  free(prev);
  free(work_d);
  free(work_u);
  return converged ? iterations : -iterations;
}

typedef struct {
  void* space;
  void* factors;
  double shift;
  int hermitian, threads;
} fs_ctx;
static void fs_hv_cb(void* c, const double* x, double* y) {
  fs_ctx* k = (fs_ctx*)c;
  long dim = ((fs_space*)k->space)->dim;
  memset(y, 0, sizeof(double) * dim);
  or_fs_hv_smp(k->space, k->factors, x, y, k->shift, k->hermitian, k->threads);
}
int or_fs_lanczos(void* space, void* factors, double shift, int hermitian, int nbrEigen, int maxIter,
                  double precision, uint64_t seed, int nbrThreads, double* eigenvalues, double* eigenvectors) {
  (void)eigenvectors;
  fs_ctx k = {space, factors, shift, hermitian, nbrThreads};
  return or_lanczos_core(fs_hv_cb, &k, ((fs_space*)space)->dim, nbrEigen, maxIter, precision, seed, eigenvalues);
}

typedef struct {
  void* space;
  const double *J, *Jz;
  int periodic, threads;
} spin_ctx;
static void spin_hv_cb(void* c, const double* x, double* y) {
  spin_ctx* k = (spin_ctx*)c;
  long dim = ((spin_space*)k->space)->dim;
  memset(y, 0, sizeof(double) * dim);
  or_spin_hv_smp(k->space, k->J, k->Jz, k->periodic, x, y, k->threads);
}
int or_spin_lanczos(void* space, const double* J, const double* Jz, int periodic, int nbrEigen, int maxIter,
                    double precision, uint64_t seed, int nbrThreads, double* eigenvalues) {
  spin_ctx k = {space, J, Jz, periodic, nbrThreads};
  return or_lanczos_core(spin_hv_cb, &k, ((spin_space*)space)->dim, nbrEigen, maxIter, precision, seed, eigenvalues);
}
