/*
 * dh_oracle.h -- CPU restatement of DiagHam's on-the-fly H.v + Lanczos path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (libdhgpu.so / libdiagham_host.so) never links or calls it.
 *
 * Every function restates the reference algorithm it cites (paths relative to
 * the DiagHam source tree).  Parity pins are listed in the .c file header.
 */
#ifndef DH_ORACLE_H
#define DH_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- fermions on the sphere ---------------- */
long or_fs_dimension(int nbrFermions, int lzMax, int totalLz);
void* or_fs_create(int nbrFermions, int totalLz, int lzMax, unsigned long memory);
void or_fs_destroy(void* space);
long or_fs_dim(void* space);
const uint64_t* or_fs_states(void* space);
const int* or_fs_state_lzmax(void* space);
long or_fs_find_state_index(void* space, uint64_t state, int lzmax);
long or_fs_adadaa(void* space, long index, int m1, int m2, int n1, int n2, double* coef);
double or_fs_aa(void* space, long index, int n1, int n2);
long or_fs_adad(void* space, int m1, int m2, double* coef);
/* n-body operators (FermionOnSphere.cc:716 ProdA, :875 ProdAd) and the n-body product with
 * explicit per-sum factor matrices (AbstractQHEOnSphereNBodyInteractionHamiltonian.h:259) */
double or_fs_prod_a(void* space, long index, const int* n, int nbrIndices);
long or_fs_prod_ad(void* space, const int* m, int nbrIndices, double* coef);
void or_nbody_hv(void* space, int nbody, int nbrSums, const int* sumCount, const int* tuples, const double* factors,
                 const double* x, double* y);

/* ---------------- Clebsch-Gordan / pseudopotentials / factors ---------------- */
void* or_cg_create(int j1, int j2);
void or_cg_destroy(void* cg);
double or_cg_get(void* cg, int m1, int m2, int j);
void or_coulomb_pseudopotentials(int nbrFlux, double* out /* nbrFlux+1 */);
void or_coulomb_closed_form(int nbrFlux, double* out /* nbrFlux+1 */);
void* or_factors_create(int lzMax, const double* pseudoPotentials /* V_0..V_lzMax */);
void or_factors_hermitian(void* factors);
void or_factors_destroy(void* factors);
int or_factors_nbr_m12(void* factors);
int or_factors_nbr(void* factors);
void or_factors_get(void* factors, int* m1, int* m2, int* nbrM3, int* m3, double* f);

/* ---------------- H.v on the sphere ---------------- */
void or_fs_hv(void* space, void* factors, const double* x, double* y, double shift, int hermitian);
void or_fs_hv_smp(void* space, void* factors, const double* x, double* y, double shift,
                  int hermitian, int nbrThreads);
long or_fs_row_elements(void* space, void* factors, long row, long* targets, double* coefs, long maxOut);
long or_fs_count_elements(void* space, void* factors);
/* every row's elements stored once (CSR) and its gather product y = H x, for solvers that apply
 * the same H many times */
void* or_fs_csr_create(void* space, void* factors, int nbrThreads);
void or_csr_destroy(void* csr);
long or_csr_nnz(void* csr);
void or_csr_hv(void* csr, const double* x, double* y, int nbrThreads);
/* rows [first, last) of (H + shift) X for k vectors ([row][k]) in gather form, into
 * out[0 .. (last - first) k) */
void or_fs_hv_rows(void* space, void* factors, const double* x, double* out, double shift, long first, long last,
                   int k, int nbrThreads);

/* ---------------- fermions in momentum space (Chern band) ---------------- */
void* or_fm_create(int nbrFermions, int nx, int ny, int kx, int ky, unsigned long memory); /* or_fs_* accessors apply */
void or_checkerboard_lower_band(int nx, int ny, double t1, double t2, double tpp, double mus, double gx, double gy,
                                double* out_u, double* out_e);
void* or_fm_factors_create(int nx, int ny, const double* u, double uPotential, double vPotential, int flatBand);
void or_fm_factors_destroy(void* factors);
int or_fm_factors_sector_size(void* factors, int sum);
void or_fm_factors_get(void* factors, int sum, int* pairs, double* fre, double* fim);
void or_fm_hv(void* space, void* factors, const double* x, double* y, int hermitian);
long or_fm_row_elements(void* space, void* factors, long row, long* targets, double* cre, double* cim, long maxOut);
/* rows [first, last) of H X for k complex vectors ([row][k][re, im]) in gather form, into
 * out[0 .. 2 (last - first) k) */
void or_fm_hv_rows(void* space, void* factors, const double* x, double* out, long first, long last, int k,
                   int nbrThreads);
/* torus two-body factors (ParticleOnTorusGenericHamiltonian with pseudopotentials when
 * coulombLevel < 0, ParticleOnTorusCoulombHamiltonian in Landau level coulombLevel otherwise) in
 * the momentum-space factor layout with Nx = 1; the torus basis of Ky is or_fm_create(N, 1, Nphi,
 * 0, Ky).  Free with or_fm_factors_destroy. */
void* or_torus_factors_create(int nflux, double ratio, int nbrPseudo, const double* pseudo, int coulombLevel);
/* FermionOnTorusWithMagneticTranslations(N, Nphi, Kx, Ky): canonical representatives ascending,
 * orbit sizes, reordering-sign bits; or_ftt_hv applies the torus factors with translation phases
 * (complex interleaved vectors). */
void* or_ftt_create(int nbrFermions, int nflux, int kx, int ky);
void or_ftt_destroy(void* space);
long or_ftt_dim(void* space);
uint64_t* or_ftt_states(void* space);
int* or_ftt_orbits(void* space);
uint64_t* or_ftt_reorder(void* space);
int or_ftt_modulo(void* space);
void or_ftt_hv(void* space, void* factors, const double* x, double* y);

/* ---------------- spin-1/2 chain ---------------- */
void* or_spin_create(int chainLength, int sz, int memorySize);
void or_spin_destroy(void* space);
long or_spin_dim(void* space);
const uint64_t* or_spin_states(void* space);
long or_spin_find_state_index(void* space, uint64_t state);
void or_spin_hv(void* space, const double* j, const double* jz, int periodic,
                const double* x, double* y);
void or_spin_hv_smp(void* space, const double* j, const double* jz, int periodic,
                    const double* x, double* y, int nbrThreads);
void or_spin_hv_rows(void* space, const double* j, const double* jz, int periodic,
                     const double* x, double* y, long first, long last, int nbrThreads);

/* ---------------- Lanczos ---------------- */
int or_tridiag_eigenvalues(int n, double* diag, double* upper, int maxIter);
int or_fs_lanczos(void* space, void* factors, double shift, int hermitian, int nbrEigen,
                  int maxIter, double precision, uint64_t seed, int nbrThreads,
                  double* eigenvalues, double* eigenvectors /* nullable: nbrEigen*dim */);
int or_spin_lanczos(void* space, const double* j, const double* jz, int periodic, int nbrEigen,
                    int maxIter, double precision, uint64_t seed, int nbrThreads, double* eigenvalues);

/* ---------------- seeded inputs ---------------- */
void or_splitmix_fill(uint64_t seed, double* out, long n);

#ifdef __cplusplus
}
#endif
#endif
