/*
 * diagham_host.h -- C entry points of the C++ host layer (libdiagham_host.so).
 *
 * The host layer mirrors DiagHam's classes on the Lanczos / H.v path (FermionOnSphere,
 * Spin1_2Chain, ParticleOnSphereGenericHamiltonian, SpinChainHamiltonian, the
 * full-reorthogonalisation Lanczos algorithms, a GPU architecture) and drives the GPU
 * through include/dhgpu.h.  These C functions expose that layer to drivers that are not
 * C++ (test harnesses, benchmarks).  Return codes follow dhgpu.h; messages of failures
 * raised by the host layer are available from dhh_last_error().
 */
#ifndef DIAGHAM_HOST_H
#define DIAGHAM_HOST_H

#include <stdint.h>

#include "dhgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dhh_arch dhh_arch;       /* GPUArchitecture */
typedef struct dhh_problem dhh_problem; /* Hilbert space + Hamiltonian of one sector */

const char* dhh_last_error(void);

int dhh_arch_create(dhh_arch** arch, int device);
int dhh_arch_destroy(dhh_arch* arch);
dh_context* dhh_arch_context(dhh_arch* arch);

/* FermionOnSphere(nbrFermions, totalLz, lzMax) + ParticleOnSphereGenericHamiltonian with
 * pseudopotentials V_0..V_lzMax (FQHESphereFermionsTwoBodyGeneric.cc:278-306) */
int dhh_sphere_problem_create(dhh_arch* arch, dhh_problem** problem, int nbr_fermions, int total_lz, int lz_max,
                              const double* pseudopotentials);
/* FermionOnSphere(nbrFermions, totalLz, lzMax) + l2_factor * L^2 (ParticleOnSphereL2Hamiltonian,
 * FQHE/src/Hamiltonian/ParticleOnSphereL2Hamiltonian.cc:240), plus the pseudopotential
 * interaction if pseudopotentials is not null (the --l2-factor option of
 * FQHESphereFermionsTwoBodyGeneric.cc).  L^2 is one more two-body term of the product. */
int dhh_sphere_l2_problem_create(dhh_arch* arch, dhh_problem** problem, int nbr_fermions, int total_lz, int lz_max,
                                 const double* pseudopotentials, double l2_factor);
/* factors of l2_factor * L^2 in the dh_sphere_hamiltonian_create layout (*nbr_factors: capacity in,
 * count out) and the constant l2_factor * Lz (Lz + 1) of the sector (*shift) */
int dhh_sphere_l2_factors(int lz_max, int nbr_fermions, int total_lz, double l2_factor, double* factors,
                          int* pair_offset, int* nbr_factors, double* shift);
/* FermionOnSphere(nbrFermions, totalLz, lzMax) + a three-body term with explicit MN factors
 * (AbstractQHEOnSphereNBodyInteractionHamiltonian.h:259-283); tuples and factors in the layout
 * of dh_sphere_nbody_hamiltonian_create (nbody = 3) */
int dhh_sphere_nbody_problem_create(dhh_arch* arch, dhh_problem** problem, int nbr_fermions, int total_lz, int lz_max,
                                    int nbody, const int* sum_count, int nbr_sums, const int* tuples,
                                    const double* factors);
/* Spin1_2Chain(L, sz) + SpinChainHamiltonian(j, jz, periodic) (GenericOpenSpinChain.cc:333-355);
 * j, jz: L entries */
int dhh_spin_problem_create(dhh_arch* arch, dhh_problem** problem, int chain_length, int sz, const double* j,
                            const double* jz, int periodic);
int dhh_problem_destroy(dhh_problem* problem);
dh_basis* dhh_problem_basis(dhh_problem* problem);
dh_hamiltonian* dhh_problem_hamiltonian(dhh_problem* problem);
/* relative angular momenta m of the adjoint channels (sphere problems); returns count */
int dhh_problem_channels(dhh_problem* problem, int* momenta, int capacity);
/* MultipleVectorHamiltonianMultiplyOperation over nbr_vectors separate device vectors of the
 * problem's dimension (src/Architecture/ArchitectureOperation/MultipleVectorHamiltonianMultiplyOperation.cc:427
 * -> AbstractHamiltonian::LowLevelMultiple[Add]Multiply, src/Hamiltonian/AbstractHamiltonian.h:344):
 * y_v = (add ? y_v : 0) + H x_v on rows [first, first + count) (count = -1: all rows). */
int dhh_problem_multiple_multiply(dhh_problem* problem, const double* const* d_x, double* const* d_y,
                                  int nbr_vectors, int add, int64_t first, int64_t count);

/* Full-reorthogonalisation (block) Lanczos to the nbr_eigen lowest eigenvalues
 * (FullReorthogonalizedLanczosAlgorithm / FullReorthogonalizedBlockLanczosAlgorithm driven
 * as in QHEOnSphereMainTask.cc:728-808).  block_size 1 = plain.  Converged when every wanted
 * eigenvalue changes by less than precision * |value| between checks and, if residual > 0,
 * every wanted Ritz pair has |H v - theta v| < residual * |theta|.  d_eigenvectors (device,
 * nbr_vectors * dim doubles, vector-major) may be null.  Outputs: eigenvalues (nbr_eigen),
 * vector products used, block steps, wall seconds (device work included). */
int dhh_lanczos(dhh_problem* problem, int nbr_eigen, int block_size, int max_iter, double precision,
                double residual, uint64_t seed, double* eigenvalues, int* products, int* blocks, double* seconds,
                double* d_eigenvectors, int nbr_vectors);

/* dhh_lanczos with disk storage (FullReorthogonalizedLanczosAlgorithmWithDiskStorage, the --disk
 * and --resume options of GenericRealMainTask.cc:120-121 and :715).  After every step each new
 * Krylov vector is written to directory/vector.<i> (DiagHam vector files, column p of block j is
 * vector j*block_size+p) and the recurrence state to directory/lanczos.dat.  resume != 0 reloads
 * both instead of starting from a random block (seed unused) and continues; the run then repeats
 * the arithmetic of an uninterrupted one, so max_iter can be raised between the two calls.
 * Errors: DH_E_ARG with "Hamiltonian does not match dimension of stored vectors in resuming" when
 * the stored vectors belong to another space, a missing or truncated file, or another block size. */
int dhh_lanczos_disk(dhh_problem* problem, const char* directory, int resume, int nbr_eigen, int block_size,
                     int max_iter, double precision, double residual, uint64_t seed, double* eigenvalues,
                     int* products, int* blocks, double* seconds, double* d_eigenvectors, int nbr_vectors);

/* Autotuner of the product launch: for each block size ks[i] it times dh_hv (after a warm-up)
 * over lanes_per_row in {1,2,4,8} (64/lanes rows per wavefront), workgroup sizes {128,256,512,1024}
 * and both index lookups, then leaves the problem configured with the fastest launch of the
 * block size best_k that maximises matrix elements * vectors per second.  Per block size the
 * table row is: k, lanes_per_row, block_size, index_mode, milliseconds per product, element-
 * vector products per second (6 doubles per row, nbr_ks rows). */
int dhh_autotune(dhh_problem* problem, const int* ks, int nbr_ks, int repeats, double* table, int* best_k);

/* Every sector total_lz[0..nbr_sectors) of nbr_fermions fermions at lz_max with the given
 * pseudopotentials, nbr_eigen lowest eigenvalues each: the Lz loop of
 * FQHESphereFermionsTwoBodyGeneric.cc:270 with a (block) full-reorthogonalisation Lanczos
 * per sector.  batched != 0 advances all sectors together (one persistent product launch
 * and segmented vector operations per step), in waves of sectors, largest first, whose expected
 * Krylov storage fits the device's free memory and memory_gb (0: no limit but the device's);
 * batched == 0 solves the sectors one after the other like the reference.  Outputs per
 * sector: eigenvalues (nbr_sectors * nbr_eigen), vector products, converged flags; total
 * wall seconds. */
/* Launch shape of the products of the batched sector solves that this thread runs next
 * (dh_sector_set_launch: lanes per row in {1, 2, 4, 8} and workgroup size; dh_sector_set_row_mode:
 * DH_INDEX_RANK, DH_INDEX_RANK_FREE or DH_INDEX_RANK_RANGE, sphere sets only); lanes_per_row = 0
 * restores the solvers' built-in choice.  No counterpart in the reference. */
int dhh_sectors_launch(int lanes_per_row, int block_size, int row_mode);

/* Batched sphere sector solves that this thread runs next: sectors of at least `rows` states are
 * solved one at a time by the single-space solver (they fill the device alone, and batching them
 * costs cache locality), the smaller ones in waves; 0 batches every sector.  Default 2,000,000.
 * memory_gb bounds the waves only: a sector solved alone grows its Krylov storage within the
 * device's free memory, as the single-space solver does.  No counterpart in the reference. */
int dhh_sectors_alone_rows(int64_t rows);

/* Autotuner of the batched sphere sector solve (no counterpart in the reference).  Calibration
 * sectors: every round(1 / fraction)-th sector by size among those with more than 256 states.  For
 * every block size ks[i]: the launch shape (lanes per row in {1, 2, 4} x workgroup in {256, 1024} x
 * row routine) with the fastest product over the calibration sectors, then a batched solve of those
 * sectors to nbr_eigen eigenvalues with that shape.  The block size whose calibration solve is
 * shortest wins: best = {k, lanes per row, workgroup, row mode}.  table: 7 doubles per ks[i]: k,
 * lanes per row, workgroup, row mode, product ms, calibration seconds (-1: the vectors did not fit),
 * vector products of the calibration solve.  Apply the result with dhh_sectors_launch. */
int dhh_sectors_autotune(dhh_arch* arch, int nbr_fermions, int lz_max, const int* total_lz, int nbr_sectors,
                         const double* pseudopotentials, int nbr_eigen, const int* ks, int nbr_ks, double fraction,
                         int max_iter, uint64_t seed, double* table, int* best);

/* Waves (sector subsets whose Krylov storage fits the device together) of the last batched sector
 * solve of this thread; 0 after a sequential solve.  No counterpart in the reference. */
int dhh_last_solve_waves(void);

int dhh_sectors_solve(dhh_arch* arch, int nbr_fermions, int lz_max, const int* total_lz, int nbr_sectors,
                      const double* pseudopotentials, int nbr_eigen, int block_size, int max_iter, double precision,
                      uint64_t seed, int batched, double memory_gb, double* eigenvalues, int* products,
                      int* converged, double* seconds);

/* FCICheckerboardLatticeModel --single-band (FTI/src/Programs/FCI/FCICheckerboardLatticeModel.cc:395-635):
 * nbr_fermions fermions in the lower band of the checkerboard model (t1, t2, tpp, mus) on an
 * Nx x Ny lattice with nearest-neighbour repulsion u_potential (and v_potential), projected;
 * momentum sectors (kx[i], ky[i]), nbr_eigen lowest eigenvalues each, complex block Lanczos with
 * block_size complex vectors, all sectors advanced together. */
int dhh_lattice_sectors_solve(dhh_arch* arch, int nbr_fermions, int nbr_site_x, int nbr_site_y, double t1,
                              double t2, double tpp, double mus, double u_potential, double v_potential,
                              int flat_band, const int* kx, const int* ky, int nbr_sectors, int nbr_eigen,
                              int block_size, int max_iter, double precision, uint64_t seed, double memory_gb,
                              double* eigenvalues, int* products, int* converged, double* seconds);
/* Host tables of that model: lower band (u: 2 complex per momentum, interleaved re/im, and
 * energies: 2 per momentum) and the two-body factors in the dh_sector_set_lattice_factors layout
 * (capacities in, counts out). */
int dhh_checkerboard_lower_band(int nbr_site_x, int nbr_site_y, double t1, double t2, double tpp, double mus,
                                double gamma_x, double gamma_y, double* u, double* energies);
int dhh_checkerboard_factors(int nbr_site_x, int nbr_site_y, double t1, double t2, double tpp, double mus,
                             double u_potential, double v_potential, int flat_band, int* sum_pair_count,
                             int* sum_pairs, int* nbr_pairs, double* factors, int64_t* nbr_factors);

/* Fermions on the torus, one sector per Ky (FermionOnTorus(N, Nphi, Ky),
 * FQHE/src/HilbertSpace/FermionOnTorus.cc:106), as in FQHETorusFermionsTwoBodyGeneric /
 * FQHETorusFermionsCoulomb: ParticleOnTorusGenericHamiltonian (nbr_pseudo pseudopotentials V_0..,
 * FQHE/src/Hamiltonian/ParticleOnTorusGenericHamiltonian.cc:244) when coulomb_level < 0, else
 * ParticleOnTorusCoulombHamiltonian in that Landau level (ParticleOnTorusCoulombHamiltonian.cc:138),
 * aspect ratio `ratio`.  The sectors run as a momentum sector set with Nx = 1 (complex vectors,
 * real factors) through the same batched complex block Lanczos as dhh_lattice_sectors_solve. */
int dhh_torus_sectors_solve(dhh_arch* arch, int nbr_fermions, int nbr_flux, double ratio, const double* pseudo,
                            int nbr_pseudo, int coulomb_level, const int* ky, int nbr_sectors, int nbr_eigen,
                            int block_size, int max_iter, double precision, uint64_t seed, double memory_gb,
                            double* eigenvalues, int* products, int* converged, double* seconds);
/* The same Hamiltonians with magnetic translations (FQHETorusFermionsWithTranslations,
 * FQHE/src/Programs/FQHEOnTorus/FQHETorusFermionsWithTranslations.cc:419): sectors (kx[i], ky[i])
 * of FermionOnTorusWithMagneticTranslations, 0 <= kx < gcd(N, Nphi), 0 <= ky < Nphi. */
int dhh_torus_translation_sectors_solve(dhh_arch* arch, int nbr_fermions, int nbr_flux, double ratio,
                                        const double* pseudo, int nbr_pseudo, int coulomb_level, const int* kx,
                                        const int* ky, int nbr_sectors, int nbr_eigen, int block_size, int max_iter,
                                        double precision, uint64_t seed, double memory_gb, double* eigenvalues,
                                        int* products, int* converged, double* seconds);
/* Its factors in the dh_sector_set_lattice_factors layout (Nx = 1, Ny = nbr_flux; capacities in,
 * counts out). */
int dhh_torus_factors(int nbr_flux, double ratio, const double* pseudo, int nbr_pseudo, int coulomb_level,
                      int* sum_pair_count, int* sum_pairs, int* nbr_pairs, double* factors, int64_t* nbr_factors);

/* DiagHam binary vector files: RealVector::WriteVector (src/Vector/RealVector.cc:3669),
 * ComplexVector::WriteVector (src/Vector/ComplexVector.cc:2606) and RealVector::ReadVector (:3798):
 * int32 dimension (-1 then int64 if it needs more than 31 bits), then little-endian doubles
 * (complex: interleaved re, im).  d_vec is device memory of dim (complex: 2 dim) doubles; reading
 * checks the file's dimension and size against dim. */
int dhh_write_vector(dhh_arch* arch, const char* file_name, const double* d_vec, int64_t dim, int complex);
int dhh_read_vector(dhh_arch* arch, const char* file_name, double* d_vec, int64_t dim, int complex);
int dhh_vector_file_dimension(const char* file_name, int complex, int64_t* dim);

/* Host-side tables (no GPU needed) */
/* LLL Coulomb pseudopotentials V_0..V_nbr_flux in CoulombPseudopotentials units; round != 0
 * rounds to the 14 significant digits written to pseudopotential files */
int dhh_coulomb_pseudopotentials(int nbr_flux, int round, double* out);
double dhh_clebsch_gordan(int j1, int j2, int m1, int m2, int j);
/* factors / pair offsets in the dh_sphere_hamiltonian_create layout; *nbr_factors is the
 * capacity on input and the count on output */
int dhh_sphere_factors(int lz_max, const double* pseudopotentials, double* factors, int* pair_offset,
                       int* nbr_factors);
/* per-pair, per-channel Clebsch-Gordan table of the adjoint split; returns channel count */
int dhh_sphere_channel_table(int lz_max, int* momenta, double* table, int capacity);
/* "Pseudopotentials = ..." file reader (FQHESpherePseudopotentialTools.cc:53) */
int dhh_read_pseudopotentials(const char* file_name, int lz_max, double* out);
/* eigenvalues of a symmetric band matrix (dense row-major n x n, half-bandwidth b) */
int dhh_band_eigenvalues(int n, int b, const double* a, double* eigenvalues);

#ifdef __cplusplus
}
#endif
#endif
