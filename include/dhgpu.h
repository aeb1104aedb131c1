/*
 * dhgpu.h -- C ABI of the MI355X (gfx950) on-the-fly Hamiltonian-vector product and
 * the Lanczos kernels behind it.
 *
 * Conventions
 *   - Every entry point returns 0 on success or a negative DH_E* code; the message of
 *     the last failure on a context is in dh_last_error(ctx).  Nothing throws across
 *     the ABI.
 *   - Pointers named d_* are device pointers (hipMalloc / dh_malloc, or a torch tensor's
 *     data_ptr()); all other pointers are host memory.
 *   - Work is enqueued on the context's stream (dh_set_stream); results are ready after
 *     dh_synchronize(ctx) (or a hipStreamSynchronize of that stream).
 *   - One context must be used from one host thread at a time.
 *   - Block vectors are interleaved [row][k]: component (row, v) of a block of k vectors
 *     is at d_x[row * k + v].  k = 1 is a plain vector.
 *   - Real FP64 throughout; complex vectors are interleaved (re, im) pairs.
 *
 * Reference interfaces replaced (paths relative to the DiagHam source tree) are cited
 * per entry point.
 */
#ifndef DHGPU_H
#define DHGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DH_OK 0
#define DH_E_ARG -1      /* invalid argument (null pointer, size, mismatched handles) */
#define DH_E_HIP -2      /* a HIP runtime call failed */
#define DH_E_NOMEM -3    /* device allocation failed */
#define DH_E_STATE -4    /* handle in the wrong state (e.g. H.v before factors are set) */
#define DH_E_RANGE -5    /* size beyond a documented limit */

typedef struct dh_context dh_context;
typedef struct dh_basis dh_basis;
typedef struct dh_hamiltonian dh_hamiltonian;

/* index lookup used by the H.v kernels */
#define DH_INDEX_BSEARCH 0 /* top-bit LUT staged in LDS, then binary search of the sorted basis */
#define DH_INDEX_RANK 1    /* combinatorial rank: two table lookups, no basis probes; on the sphere
                              the two-body created pairs come from the free-pair masks for k <= 2
                              and from the range loop for k > 2, the faster of the two per k */
#define DH_INDEX_RANK_X4 2 /* rank, with the two-body created-pair loop unrolled four-wide (sphere only;
                              other Hamiltonians treat it as DH_INDEX_RANK) */
#define DH_INDEX_RANK_FREE 3 /* rank, with the created pairs taken from the mask of free pairs of each
                                annihilated pair for every k (sphere only; other Hamiltonians treat it
                                as DH_INDEX_RANK) */
#define DH_INDEX_RANK_RANGE 4 /* rank, with every created pair of the pair-sum range visited and the
                                 occupied ones skipped, for every k (sphere only; other Hamiltonians
                                 treat it as DH_INDEX_RANK) */

/* ---------------------------------------------------------------- context
 * Replaces the per-operation thread fork/join of SMPArchitecture::SendJobs
 * (src/Architecture/SMPArchitecture.cc:134): one context owns a device and a stream. */
int dh_create(dh_context** ctx, int device);
int dh_destroy(dh_context* ctx);
const char* dh_last_error(dh_context* ctx);
int dh_set_stream(dh_context* ctx, void* hip_stream);
int dh_synchronize(dh_context* ctx);
/* number of compute units and gfx name of the context's device (name >= 64 bytes) */
int dh_device_info(dh_context* ctx, int* nbr_cu, char* name);
/* free and total device memory of the context's device (hipMemGetInfo); no counterpart in the
 * reference: the host sizes the Krylov storage it grows on demand by it */
int dh_memory_info(dh_context* ctx, size_t* free_bytes, size_t* total_bytes);
/* Highest device memory in use (the whole device's, sampled after every allocation this context
 * makes) since the context was created or the last reset; reset != 0 starts a new period.  No
 * counterpart in the reference: it reports the footprint of a solve. */
int dh_memory_peak(dh_context* ctx, size_t* peak_bytes, int reset);

/* Device vector storage; AbstractArchitecture::GetNewRealVector / AllocateMemory hooks
 * (src/Architecture/AbstractArchitecture.h:150-246). */
int dh_malloc(dh_context* ctx, void** d_ptr, size_t bytes);
int dh_free(dh_context* ctx, void* d_ptr);
int dh_memcpy_h2d(dh_context* ctx, void* d_dst, const void* src, size_t bytes);
int dh_memcpy_d2h(dh_context* ctx, void* dst, const void* d_src, size_t bytes);
int dh_memset(dh_context* ctx, void* d_dst, int value, size_t bytes);

/* ---------------------------------------------------------------- bases
 * FermionOnSphere(nbrFermions, totalLz, lzMax) constructor + GenerateStates +
 * GenerateLookUpTable (FQHE/src/HilbertSpace/FermionOnSphere.cc:76, :1621, :1659).
 * The basis is generated on the device by unranking in the reference's order (states
 * sorted as descending unsigned 64-bit integers, bit i = orbital l_z = i - lzMax/2).
 * lzMax <= 62. */
int dh_fermion_sphere_create(dh_context* ctx, dh_basis** basis, int nbr_fermions, int total_lz, int lz_max);
/* EvaluateHilbertSpaceDimension (FermionOnSphere.cc:1794); no device work. */
int64_t dh_fermion_sphere_dimension(int nbr_fermions, int total_lz, int lz_max);
/* A basis from a StateDescription that DiagHam already built (FermionOnSphere::StateDescription,
 * FermionOnSphere.cc:1621: dim host words, descending).  The device lookups index the full
 * sector, so the words must be exactly that sector in the reference's order; they are checked
 * against the device's own generation (DH_E_ARG, with the first differing index in
 * dh_last_error, when they are not). */
int dh_fermion_sphere_from_states(dh_context* ctx, dh_basis** basis, int nbr_fermions, int total_lz, int lz_max,
                                  const uint64_t* states, int64_t dim);

/* Spin1_2Chain(chainLength, sz, memory) (Spin/src/HilbertSpace/Spin1_2Chain.cc:113, :331):
 * states ordered as the reference (site 0 most significant, bits >= L kept at 1 in the
 * 32-bit word).  chain_length <= 32. */
int dh_spin_chain_create(dh_context* ctx, dh_basis** basis, int chain_length, int sz);

int dh_basis_destroy(dh_basis* basis);
int64_t dh_basis_dimension(const dh_basis* basis);
/* copy the state words (StateDescription) to the host: 8 bytes per state */
int dh_basis_get_states(dh_basis* basis, uint64_t* states);
/* FindStateIndex (FermionOnSphere.cc:1431, Spin1_2Chain.cc:1002) for n states at once,
 * evaluated by the device index lookup `mode`; out-of-space states give dimension. */
int dh_basis_find_state_index(dh_basis* basis, const uint64_t* states, int64_t n, int mode, int64_t* indices);

/* ---------------------------------------------------------------- Hamiltonians
 * ParticleOnSphereGenericHamiltonian (FQHE/src/Hamiltonian/ParticleOnSphereGenericHamiltonian.cc:68)
 * with the full (non-Hermitian-reduced) two-body factor list of EvaluateInteractionFactors
 * (:493): for every annihilated pair m1 > m2 (pair index m1(m1-1)/2 + m2) the factors of
 * the created pairs m3 > m4 = m1 + m2 - m3, m3 ascending, are factors[pair_offset[pair] +
 * (m3 - (m1 + m2) / 2 - 1)].  nbr_factors = pair_offset[nbrPairs].  The host computes the
 * factors; this call copies them to the device.  The factors of a Hermitian two-body term are
 * symmetric, F(a -> b) = F(b -> a); the product kernels keep one triangle per pair sum, and a
 * table that is not symmetric to 1e-12 is refused with DH_E_ARG. */
int dh_sphere_hamiltonian_create(dh_context* ctx, dh_hamiltonian** ham, dh_basis* basis,
                                 const double* factors, const int* pair_offset, int nbr_factors);
/* Three-body term of AbstractQHEOnSphereNBodyInteractionHamiltonian with explicit MN factors
 * (FQHE/src/Hamiltonian/AbstractQHEOnSphereNBodyInteractionHamiltonian.h:259-283, the
 * MNNBodyInteractionFactors branch; FermionOnSphere::ProdA :716 / ProdAd :875):
 *   H = sum_S sum_{i,l} F_S[i][l] ProdAd(M_l) ProdA(N_i).
 * nbody must be 3.  Orbital sum S = 0 .. nbr_sums-1 lists sum_count[S] tuples of three orbitals in
 * descending order (as GetAllSkewSymmetricIndices :1600 orders them), the sums' tuples one after
 * the other in `tuples`; F_S (sum_count[S]^2 doubles, row i = annihilated tuple i) follow each
 * other in `factors`.  F_S must be symmetric (the product is gathered per destination row). */
int dh_sphere_nbody_hamiltonian_create(dh_context* ctx, dh_hamiltonian** ham, dh_basis* basis, int nbody,
                                       const int* sum_count, int nbr_sums, const int* tuples,
                                       const double* factors);
/* Per-channel factor split used by the adjoint kernel: the two-body factor of (pair a ->
 * pair b) is sum_c V_c * (-4) * cg[a * nbr_channels + c] * cg[b * nbr_channels + c], with
 * V_c = channel_weights[c]. */
int dh_sphere_hamiltonian_set_channels(dh_hamiltonian* ham, const double* cg, int nbr_channels,
                                       const double* channel_weights);

/* SpinChainHamiltonian(chain, nbrSpin, j, jz, periodic) (Spin/src/Hamiltonian/SpinChainHamiltonian.cc:97):
 * j and jz have chain_length entries (bond i couples sites i and i+1 mod L). */
int dh_spin_chain_hamiltonian_create(dh_context* ctx, dh_hamiltonian** ham, dh_basis* basis,
                                     const double* j, const double* jz, int periodic);

int dh_hamiltonian_destroy(dh_hamiltonian* ham);
/* AbstractHamiltonian::ShiftHamiltonian (src/Hamiltonian/AbstractHamiltonian.h) */
int dh_hamiltonian_set_shift(dh_hamiltonian* ham, double shift);
int64_t dh_hamiltonian_dimension(const dh_hamiltonian* ham);

/* Launch shape of the H.v kernels.  lanes_per_row in {1,2,4,8} (64/lanes_per_row destination
 * rows per wavefront; spin chains: 1); block_size a multiple of 64 up to 1024; index_mode
 * DH_INDEX_*; lds_tables != 0 stages the index tables in LDS when they fit.  Until it is called, a
 * sphere Hamiltonian picks its shape per product: on spaces of at least 256 rows per CU the
 * fastest shapes measured on N=12 2S=33 Lz=0 (with the 16-bit low rank table in LDS: one lane per
 * row, 1024-thread workgroups for k = 1, two lanes per row for k = 2, 512 threads for k > 2; with the
 * rank tables in global memory:
 * two lanes per row, 256 threads for k <= 2 and 1024 above); on smaller spaces four lanes per row,
 * 512 threads. */
int dh_hamiltonian_set_launch(dh_hamiltonian* ham, int lanes_per_row, int block_size, int index_mode, int lds_tables);

/* ---------------------------------------------------------------- products
 * y = (add ? y : 0) + (H + shift) x for a block of k vectors ([row][k] layout).
 * LowLevelAddMultiply / LowLevelMultiply (src/Hamiltonian/AbstractHamiltonian.h:265,
 * AbstractHamiltonian.cc:468) for k = 1 and LowLevelMultipleAddMultiply
 * (AbstractHamiltonian.h:344; FQHE AbstractQHEOnSphereHamiltonian.cc:1217) for k > 1;
 * rows [first, first + count) of y are written (count = -1: all rows).
 * k in {1, 2, 4, 8, 16}.  Gather form: every destination row is written once, no atomics. */
int dh_hv(dh_hamiltonian* ham, const double* d_x, double* d_y, int k, int add, int64_t first, int64_t count);

/* LowLevelMultipleAddMultiply(RealVector* vSources, RealVector* vDestinations, int nbrVectors,
 * int firstComponent, int nbrComponent) (src/Hamiltonian/AbstractHamiltonian.h:344), as called by
 * MultipleVectorHamiltonianMultiplyOperation::RawApplyOperation
 * (src/Architecture/ArchitectureOperation/MultipleVectorHamiltonianMultiplyOperation.cc:427):
 * y_v = (add ? y_v : 0) + (H + shift) x_v for v < nbr_vectors, each vector a separate device
 * array of the full dimension (d_x, d_y: host arrays of nbr_vectors device pointers); rows
 * [first, first + count) of every y_v are written (count = -1: all).  Any nbr_vectors >= 0: groups
 * of 16, 8, 4 and 2 are interleaved into a [row][k] staging block of the context (grown on demand,
 * 16 dim k bytes, kept until the context is freed) so that they share every index lookup. */
int dh_hv_vectors(dh_hamiltonian* ham, const double* const* d_x, double* const* d_y, int nbr_vectors, int add,
                  int64_t first, int64_t count);

/* AbstractQHEOnSphereHamiltonian::EnableFastMultiplication (FQHE/src/Hamiltonian/
 * AbstractQHEOnSphereHamiltonian.cc:3814): generate every matrix element once into device
 * memory (CSR: int64 row offsets, int32 columns, FP64 values; the shift stays separate) when that
 * takes at most max_bytes; later dh_hv calls stream the stored elements.  *bytes (may be null)
 * receives the size the elements need.  If it exceeds max_bytes the call returns DH_E_NOMEM and
 * the products stay on the fly.  Two-body sphere Hamiltonians only. */
int dh_hamiltonian_precompute(dh_hamiltonian* ham, int64_t max_bytes, int64_t* bytes);
/* free the stored elements: products go back to on the fly */
int dh_hamiltonian_release_precomputed(dh_hamiltonian* ham);
/* Number of matrix elements per product (full non-Hermitian-reduced terms with a target in
 * the space, diagonal terms included), counted on the device. */
int dh_hamiltonian_count_elements(dh_hamiltonian* ham, int64_t* nnz);

/* Matrix elements of row `row` in the order the kernel visits them: (target index,
 * coefficient).  Returns the count in *n (at most max_out are written). */
int dh_hamiltonian_row_elements(dh_hamiltonian* ham, int64_t row, int64_t* targets, double* coefs,
                                int64_t max_out, int64_t* n);

/* ---------------------------------------------------------------- many sectors in one launch
 * The Lz loop of FQHESphereFermionsTwoBodyGeneric.cc:270 (one FermionOnSphere + Hamiltonian +
 * Lanczos per sector, run one after the other) becomes one set of sectors: the bases are
 * concatenated (sector s owns global rows [offsets[s], offsets[s+1])), block vectors of the set
 * are the per-sector [D_s][k] blocks concatenated, and every call below covers all sectors with
 * active[s] != 0 (active = null: all) in one persistent launch. */
typedef struct dh_sector_set dh_sector_set;
int dh_sector_set_create(dh_context* ctx, dh_sector_set** set, int nbr_fermions, int lz_max, const int* total_lz,
                         int nbr_sectors);
int dh_sector_set_destroy(dh_sector_set* set);
int64_t dh_sector_set_dimension(const dh_sector_set* set);
int dh_sector_set_offsets(const dh_sector_set* set, int64_t* offsets /* nbr_sectors + 1 */);
int dh_sector_set_get_states(dh_sector_set* set, uint64_t* states);
/* factors in the dh_sphere_hamiltonian_create layout (shared by every sector) and the shift */
int dh_sector_set_factors(dh_sector_set* set, const double* factors, const int* pair_offset, int nbr_factors,
                          double shift);
int dh_sector_set_launch(dh_sector_set* set, int lanes_per_row, int block_size);
/* Row routine of the products of a sphere or momentum set: DH_INDEX_RANK_RANGE (every created pair
 * of the range or list, skipping occupied ones), DH_INDEX_RANK_FREE (only the free created pairs,
 * from one bit mask per annihilated pair) or DH_INDEX_RANK (sphere sets: free pairs for k <= 2, the
 * range for k > 2; momentum sets: the list).  A tuning knob like dh_sector_set_launch; the results
 * agree to rounding. */
int dh_sector_set_row_mode(dh_sector_set* set, int mode);
/* Launch shape of the segmented vector operations of a set (no counterpart in the reference, a
 * tuning knob like dh_sector_set_launch): persistent axpy / rmul workgroups per CU (default 16) and
 * rows per dots tile (0: automatic). */
int dh_sector_set_blas_launch(dh_sector_set* set, int axpy_groups_per_cu, int64_t dots_rows_per_tile);
/* Matrix elements per product of each sector of a sphere or momentum set (counts[nbr_sectors]),
 * counted as dh_hamiltonian_count_elements does: the "nbr interaction" the reference prints per
 * sector (AbstractQHEOnSphereHamiltonian.cc:3721 PartialFastMultiplicationMemory); for momentum
 * sets every (annihilated pair, created pair) term of a momentum sum whose target is in the space. */
int dh_sector_set_count_elements(dh_sector_set* set, int64_t* counts);
/* y = (add ? y : 0) + (H + shift) x over the active sectors (LowLevelMultipleAddMultiply per sector) */
int dh_sectors_hv(dh_sector_set* set, const double* d_x, double* d_y, int k, int add, const unsigned char* active);
/* per active sector j (in sector order): c[j][(b*k+p)*k+q] = sum_rows basis_b[row][p] w[row][q];
 * the basis holds nbr_blocks set-blocks, block b at d_basis + b * dimension * k */
int dh_sectors_block_dots(dh_sector_set* set, const double* d_basis, int nbr_blocks, const double* d_w, int k,
                          const unsigned char* active, double* c);
/* w[row][q] += sum_{b,p} basis_b[row][p] c[j][(b*k+p)*k+q] for the rows of active sector j */
int dh_sectors_block_axpy(dh_sector_set* set, const double* d_basis, int nbr_blocks, double* d_w, int k,
                          const unsigned char* active, const double* c);
/* w[row] <- w[row] * m[j] (k x k per active sector j) */
int dh_sectors_block_rmul(dh_sector_set* set, double* d_w, int k, const unsigned char* active, const double* m);
/* w <- w m[j] + sum_b basis_b c[j][b] over the rows of active sector j (m may be null: identity),
 * then dots[j][b] = basis_b^T w (dots may be null) and gram[j] = w^T w over the new w, in one pass
 * over the basis (k = 1 or 2 doubles per row).  Returns DH_E_RANGE when nbr_blocks is too large for
 * the staged tile; the separate axpy / dots / rmul calls do the same work. */
int dh_sectors_block_update(dh_sector_set* set, const double* d_basis, int nbr_blocks, double* d_w, int k,
                            const unsigned char* active, const double* c, const double* m, double* dots, double* gram);

/* Momentum sectors (kx[i], ky[i]) of nbr_fermions spinless fermions in one band of an Nx x Ny
 * lattice: FermionOnSquareLatticeMomentumSpace (FQHE/src/HilbertSpace/FermionOnSquareLatticeMomentumSpace.cc:76),
 * orbital kx * Ny + ky, states sorted as descending 64-bit integers, Nx * Ny <= 62.  Vectors of
 * such a set are complex, interleaved (re, im): a block of k complex vectors is [row][k][2]
 * doubles, and dh_sectors_block_* take 2k as their k. */
int dh_momentum_sector_set_create(dh_context* ctx, dh_sector_set** set, int nbr_fermions, int nbr_site_x,
                                  int nbr_site_y, const int* kx, const int* ky, int nbr_sectors);
/* FermionOnSquareLatticeMomentumSpace::EvaluateHilbertSpaceDimension (FermionOnSquareLatticeMomentumSpace.cc:290):
 * the number of states of the momentum sector (kx, ky); no device work; -1 for arguments outside
 * the ranges above. */
int64_t dh_momentum_sector_dimension(int nbr_fermions, int nbr_site_x, int nbr_site_y, int kx, int ky);
/* Two-body factors of ParticleOnLatticeTimeReversalBreakingSingleBandHamiltonian
 * (FQHE/src/Hamiltonian/ParticleOnLatticeTimeReversalBreakingSingleBandHamiltonian.h:370):
 * momentum sum j = 0..nbr_sums-1 (nbr_sums = Nx*Ny) lists sum_pair_count[j] orbital pairs
 * (a < b) in sum_pairs (2 ints per pair, sums in order), and its complex factor matrix
 * F_j[i1][i2] (annihilated pair i1, created pair i2) follows in factors, (re, im) interleaved,
 * the matrices of the sums concatenated. */
int dh_sector_set_lattice_factors(dh_sector_set* set, int nbr_sums, const int* sum_pair_count, const int* sum_pairs,
                                  const double* factors, double shift);

/* Fermions on a torus of nbr_flux flux quanta with magnetic translations, sectors (kx[i], ky[i])
 * with 0 <= kx < gcd(N, Nphi), 0 <= ky < Nphi: FermionOnTorusWithMagneticTranslations(N, Nphi,
 * Kx, Ky) (FQHE/src/HilbertSpace/FermionOnTorusWithMagneticTranslations.cc:80): per sector the
 * smallest state of each orbit of the translation by Nphi / gcd(N, Nphi) orbitals, ascending, the
 * orbits allowed by Kx only.  Complex vectors as for momentum sets.  Its two-body factors go in
 * with dh_sector_set_lattice_factors (nbr_sums = Nphi, the torus layout of dhh_torus_factors); the
 * product applies AbstractQHEOnTorusWithMagneticTranslationsHamiltonian.h:160-190 (translation
 * phases exp(2 pi i Kx t / gcd), orbit rescaling and reordering signs). */
int dh_torus_translation_sector_set_create(dh_context* ctx, dh_sector_set** set, int nbr_fermions, int nbr_flux,
                                           const int* kx, const int* ky, int nbr_sectors);
/* per row of the set: orbit size (NbrStateInOrbit) and reordering-sign bits (ReorderingSign);
 * either pointer may be null */
int dh_sector_set_torus_data(dh_sector_set* set, uint8_t* orbits, uint64_t* reorder);

/* ---------------------------------------------------------------- adjoint (no reference counterpart)
 * g[c] = <psi| H_c |phi> for every channel c of dh_sphere_hamiltonian_set_channels, in one
 * pass over the basis; if d_hpsi is not null it also receives H psi (the vector-Jacobian
 * product of x -> H x for the cotangent psi).  Hellmann-Feynman companion of
 * --energy-expectation (FQHE/src/Programs/FQHEOnSphere/FQHESphereFermionsTwoBodyGeneric.cc:323). */
int dh_sphere_adjoint(dh_hamiltonian* ham, const double* d_psi, const double* d_phi, double* g, double* d_hpsi);

/* ---------------------------------------------------------------- Lanczos BLAS
 * Krylov bases are arrays of nbr_blocks blocks of k vectors, block b at d_basis + b * dim * k,
 * each block in [row][k] layout.
 * MultipleRealScalarProductOperation (src/Architecture/ArchitectureOperation/MultipleRealScalarProductOperation.cc:174):
 *   c[(b * k + p) * k + q] = sum_row basis_b[row][p] * w[row][q]   (host result, nbr_blocks*k*k) */
int dh_block_dots(dh_context* ctx, const double* d_basis, int nbr_blocks, const double* d_w, int64_t dim,
                  int k, double* c);

/* dh_block_dots without any allocation, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * The row-chunk partial sums live in a fixed device array of 2^19 doubles (4 MiB) per device, held
 * until the process ends; the result goes to device memory d_out (nbr_blocks * k * k doubles, the
 * layout of dh_block_dots) and is valid after the stream has reached it (no wait, no pinned
 * staging).  nbr_blocks * k * k above 2^19 is refused with DH_E_RANGE.  Every call on one device
 * shares that array: calls that overlap on one device race. */
int dh_block_dots_noalloc(dh_context* ctx, const double* d_basis, int nbr_blocks, const double* d_w, int64_t dim,
                          int k, double* d_out);
/* AddRealLinearCombinationOperation (AddRealLinearCombinationOperation.cc:157):
 *   w[row][q] += sum_{b,p} basis_b[row][p] * c[(b * k + p) * k + q]   (c on the host) */
int dh_block_axpy(dh_context* ctx, const double* d_basis, int nbr_blocks, double* d_w, int64_t dim, int k,
                  const double* c);
/* w[row][q] = sum_p w[row][p] * m[p * k + q]  (in-place right multiplication by a k x k host matrix) */
int dh_block_rmul(dh_context* ctx, double* d_w, int64_t dim, int k, const double* m);
/* w <- w m + sum_b basis_b c[b] (m may be null: identity), then dots[b] = basis_b^T w (dots may be
 * null) and gram = w^T w over the new w, in one pass over the basis (k = 1 or 2): the middle step of
 * the two-round Gram-Schmidt of FullReorthogonalizedLanczosAlgorithm::RunLanczosAlgorithm
 * (src/LanczosAlgorithm/FullReorthogonalizedLanczosAlgorithm.cc:240-250).  DH_E_RANGE when
 * nbr_blocks is too large for the staged tile. */
int dh_block_update(dh_context* ctx, const double* d_basis, int nbr_blocks, double* d_w, int64_t dim, int k,
                    const double* c, const double* m, double* dots, double* gram);
/* copy column `col` of a [row][k] block into / out of a plain vector of dim doubles */
int dh_block_get_column(dh_context* ctx, const double* d_w, int64_t dim, int k, int col, double* d_out);
int dh_block_set_column(dh_context* ctx, double* d_w, int64_t dim, int k, int col, const double* d_in);
/* y = x over n doubles by a streaming kernel (also the HBM-traffic calibration kernel) */
int dh_copy(dh_context* ctx, double* d_y, const double* d_x, int64_t n);
/* y = a * x + b * y over n doubles */
int dh_axpby(dh_context* ctx, double a, const double* d_x, double b, double* d_y, int64_t n);
/* uniform [-1, 1) SplitMix64 values (seeded start vectors) */
int dh_fill_random(dh_context* ctx, double* d_x, int64_t n, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif /* DHGPU_H */
