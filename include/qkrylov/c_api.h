/*
 * qkrylov C ABI — drop-in replacement for the reference `include/qkrylov/c_api.h`.
 *
 * Part 1 keeps every declaration of the reference header with the same name,
 * signature, struct layout, macro value and error convention
 * (reference: include/qkrylov/c_api.h:14-168, behaviour: src/c_api.cpp:33-626).
 * The float entry points compute in FP64 on the GPU and narrow their outputs.
 *
 * Part 2 adds FP64, device-resident and batched entry points that the float
 * ABI cannot carry (1e-12 / 1e-10 parity bars, inputs already in HBM,
 * several sectors or starting vectors in one launch).  They have no
 * counterpart in the reference; each comment names the reference routine
 * whose semantics it extends.
 *
 * Threading: a Hamiltonian handle owns one HIP stream and device buffers;
 * calls on one handle must not overlap.  Distinct handles may be used from
 * distinct threads.
 */
#ifndef QKRYLOV_C_API_H
#define QKRYLOV_C_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * Return codes (reference c_api.h:14-16)
 * ------------------------------------------------------------------------- */
#define QKRYLOV_SUCCESS                0
#define QKRYLOV_ERROR_INVALID_ARG     -1
#define QKRYLOV_ERROR_EXCEPTION       -2

/* ---------------------------------------------------------------------------
 * Opaque handles (reference c_api.h:21-25)
 * ------------------------------------------------------------------------- */
typedef struct qkrylov_sector_t*      qkrylov_sector_h;
typedef struct qkrylov_basis_t*       qkrylov_basis_h;
typedef struct qkrylov_site_t*        qkrylov_site_h;
typedef struct qkrylov_opsum_t*       qkrylov_opsum_h;
typedef struct qkrylov_hamiltonian_t* qkrylov_hamiltonian_h;

/* ===========================================================================
 * Part 1 — reference ABI
 * ========================================================================= */

/* Sector (reference c_api.h:30-39, c_api.cpp:56-158) */
qkrylov_sector_h qkrylov_sector_create(void);
void             qkrylov_sector_destroy(qkrylov_sector_h sector);
int              qkrylov_sector_set_sz(qkrylov_sector_h sector, int sz2);
int              qkrylov_sector_set_hubbard_particles(qkrylov_sector_h sector, int nup, int ndn);
int              qkrylov_sector_set_n(qkrylov_sector_h sector, int n);
int              qkrylov_sector_set_nb(qkrylov_sector_h sector, int nb);
int              qkrylov_sector_get_sz(qkrylov_sector_h sector, int* sz2_out, int* active_out);
int              qkrylov_sector_get_hubbard_particles(qkrylov_sector_h sector, int* nup_out, int* ndn_out, int* active_out);
int              qkrylov_sector_get_n(qkrylov_sector_h sector, int* n_out, int* active_out);
int              qkrylov_sector_get_nb(qkrylov_sector_h sector, int* nb_out, int* active_out);

/* Basis (reference c_api.h:44-53, c_api.cpp:161-257).  States are ordered by
 * ascending integer value, exactly as the reference enumerates them
 * (src/basis/<kind>_basis.cpp); index/state use closed-form ranking tables. */
qkrylov_basis_h  qkrylov_spinhalf_basis_create(int num_sites, qkrylov_sector_h sector);
qkrylov_basis_h  qkrylov_fermion_basis_create(int num_sites, qkrylov_sector_h sector);
qkrylov_basis_h  qkrylov_hubbard_basis_create(int num_sites, qkrylov_sector_h sector);
qkrylov_basis_h  qkrylov_tj_basis_create(int num_sites, qkrylov_sector_h sector);
void             qkrylov_basis_destroy(qkrylov_basis_h basis);
uint64_t         qkrylov_basis_dimension(qkrylov_basis_h basis);
int              qkrylov_basis_nsites(qkrylov_basis_h basis);
uint64_t         qkrylov_basis_state(qkrylov_basis_h basis, uint64_t index);
int64_t          qkrylov_basis_index(qkrylov_basis_h basis, uint64_t state_bitstring);
int              qkrylov_basis_contains(qkrylov_basis_h basis, uint64_t state_bitstring);

/* Site (reference c_api.h:58-62, c_api.cpp:260-302) */
qkrylov_site_h   qkrylov_spinhalf_site_create(void);
qkrylov_site_h   qkrylov_fermion_site_create(void);
qkrylov_site_h   qkrylov_hubbard_site_create(void);
qkrylov_site_h   qkrylov_tj_site_create(void);
void             qkrylov_site_destroy(qkrylov_site_h site);

/* OpSum (reference c_api.h:67-76, c_api.cpp:305-373).  Factors act on the ket
 * in list order; coefficients are complex. */
qkrylov_opsum_h  qkrylov_opsum_create(void);
void             qkrylov_opsum_destroy(qkrylov_opsum_h opsum);
int              qkrylov_opsum_clear(qkrylov_opsum_h opsum);
int              qkrylov_opsum_add_term_1body(qkrylov_opsum_h opsum, float coeff_real, float coeff_imag,
                                             const char* op1, int site1);
int              qkrylov_opsum_add_term_2body(qkrylov_opsum_h opsum, float coeff_real, float coeff_imag,
                                             const char* op1, int site1,
                                             const char* op2, int site2);
int              qkrylov_opsum_add_term_nbody(qkrylov_opsum_h opsum, float coeff_real, float coeff_imag,
                                             int n_factors, const char** ops, const int* sites);

/* Hamiltonian (reference c_api.h:81-100, c_api.cpp:376-471).  create() compiles
 * the OpSum into bit-operation groups and uploads ranking tables; there is no
 * CSR.  apply computes y[a] = sum_terms conj(amp(a->b)) x[b], the reference's
 * gather convention (src/hamiltonian/matrix_free_hamiltonian.cpp:45-91). */
qkrylov_hamiltonian_h qkrylov_hamiltonian_create(qkrylov_basis_h basis,
                                                qkrylov_site_h site,
                                                qkrylov_opsum_h opsum);
void                  qkrylov_hamiltonian_destroy(qkrylov_hamiltonian_h h);
uint64_t              qkrylov_hamiltonian_dimension(qkrylov_hamiltonian_h h);
int                   qkrylov_hamiltonian_apply(qkrylov_hamiltonian_h h,
                                                const float* x_real, const float* x_imag,
                                                float* y_real, float* y_imag);
/* x_complex, y_complex: 2*dimension() floats, interleaved [re, im, ...]. */
int                   qkrylov_hamiltonian_apply_complex(qkrylov_hamiltonian_h h,
                                                        const float* x_complex,
                                                        float* y_complex);
/* Real part of H_ii into diag_out[dimension()]. */
int                   qkrylov_hamiltonian_diagonal(qkrylov_hamiltonian_h h,
                                                   float* diag_out);

/* Solvers (reference c_api.h:105-168, c_api.cpp:474-624) */
typedef struct {
    float energy;
    int iterations;
    int converged;
} qkrylov_lanczos_result_c_t;

typedef struct {
    int iterations;
    int converged;
} qkrylov_davidson_result_c_t;

int qkrylov_lanczos_ground_state(qkrylov_hamiltonian_h h,
                                 int maxiter,
                                 float tol,
                                 qkrylov_lanczos_result_c_t* result);

int qkrylov_lanczos_ground_state_complex(qkrylov_hamiltonian_h h,
                                         int maxiter,
                                         float tol,
                                         qkrylov_lanczos_result_c_t* result,
                                         float* eigenvector_complex);

int qkrylov_davidson_lowest_complex(qkrylov_hamiltonian_h h,
                                    int n_eig,
                                    int max_subspace,
                                    float tol,
                                    float* eigenvalues_out,
                                    float* eigenvectors_complex_out,
                                    qkrylov_davidson_result_c_t* result_info);

int qkrylov_continued_fraction_coeffs_complex(qkrylov_hamiltonian_h h,
                                              const float* phi0_complex,
                                              int n_iter,
                                              float* alphas_out,
                                              float* betas_out,
                                              float* norm_phi0_out,
                                              int* num_coeffs_out);

float qkrylov_evaluate_spectral_function(const float* alphas,
                                         const float* betas,
                                         size_t n,
                                         float norm_phi0,
                                         float omega,
                                         float E0,
                                         float eta);

typedef struct {
    float beta;
    float partition_function;
    float internal_energy;
    float specific_heat;
} qkrylov_ftlm_result_c_t;

int qkrylov_ftlm(qkrylov_hamiltonian_h h,
                 float beta,
                 int n_random,
                 int n_steps,
                 qkrylov_ftlm_result_c_t* result);

/* ===========================================================================
 * Part 2 — FP64, device-resident and batched extensions (no reference
 * counterpart; semantics follow the reference routine named in each comment)
 * ========================================================================= */

typedef struct {
    double energy;
    int iterations;
    int converged;
    /* steps whose reorthogonalisation needed the second Gram-Schmidt pass */
    int second_passes;
} qkrylov_lanczos_result_f64_t;

typedef struct {
    double beta;
    double partition_function;
    double internal_energy;
    double specific_heat;
} qkrylov_ftlm_result_f64_t;

/* Last error message of the calling thread ("" if none). */
const char* qkrylov_last_error(void);

/* Device name and basic properties, for logs. Returns 0 on success. */
int qkrylov_device_info(char* buf, size_t len);
/* Number of visible GPUs (0 when the HIP runtime finds none). */
int qkrylov_device_count(void);

/* Sector: clear every constraint flag. */
int qkrylov_sector_reset(qkrylov_sector_h sector);

/* OpSum with FP64 coefficients (reference c_api.cpp:358-373 with Real=double). */
int qkrylov_opsum_add_term_f64(qkrylov_opsum_h opsum, double coeff_real, double coeff_imag,
                               int n_factors, const char** ops, const int* sites);
size_t qkrylov_opsum_size(qkrylov_opsum_h opsum);

/* Bulk state/index maps: states_out[i] = state(first + i) for i < count;
 * index_out[i] = index(states[i]) or -1 (reference basis state()/index()). */
int qkrylov_basis_states(qkrylov_basis_h basis, uint64_t first, uint64_t count, uint64_t* states_out);
int qkrylov_basis_indices(qkrylov_basis_h basis, const uint64_t* states, uint64_t count, int64_t* index_out);

/* y = H x with host FP64 buffers, interleaved complex, nvec vectors stored
 * one after another (vector v at offset v*2*dim doubles).  Reference:
 * MatrixFreeHamiltonian::apply(const Complex*, Complex*), cpp:174-201. */
int qkrylov_hamiltonian_apply_f64(qkrylov_hamiltonian_h h, const double* x_complex,
                                  double* y_complex, int nvec);

/* y = H x with device-resident buffers (complex FP64, vector v at offset
 * v*ld elements).  stream: a hipStream_t, or NULL for the handle's stream.
 * Asynchronous with respect to the host.  Reference: apply(View, View),
 * cpp:141-168. */
int qkrylov_hamiltonian_apply_device_f64(qkrylov_hamiltonian_h h, const void* x_dev, void* y_dev,
                                         int nvec, uint64_t ld, void* stream);

/* ---- Operator between bases.  No counterpart in the reference, which applies
 * only Hamiltonians within one basis; the operator-on-state apply its
 * dynamics tests ask for (tests/test_dynamics.cpp:43-46): y = O x for an
 * OpSum O taking states of basis_in to basis_out, which must share site kind,
 * size and sector type (two Sz sectors, two particle-number sectors, two
 * (nup, ndn) sectors); O need not be Hermitian nor conserve the sector
 * (S+(q), c_k).  Same factor semantics as the Hamiltonian (factors act on the
 * ket in list order); y[b] = sum_a <b|O|a> x[a]. ---- */
typedef struct qkrylov_operator_t* qkrylov_operator_h;
qkrylov_operator_h qkrylov_operator_create(qkrylov_basis_h basis_in, qkrylov_basis_h basis_out,
                                          qkrylov_site_h site, qkrylov_opsum_h opsum);
void qkrylov_operator_destroy(qkrylov_operator_h op);
uint64_t qkrylov_operator_dimension_in(qkrylov_operator_h op);
uint64_t qkrylov_operator_dimension_out(qkrylov_operator_h op);
/* Host FP64 buffers: nvec input vectors of dimension_in complex values one
 * after another, nvec outputs of dimension_out. */
int qkrylov_operator_apply_f64(qkrylov_operator_h op, const double* x_complex, double* y_complex,
                               int nvec);
/* Device buffers: input v at x_dev + v*ld_in, output v at y_dev + v*ld_out
 * complex FP64 elements (0: the dimension); asynchronous on stream (NULL: the
 * handle's stream). */
int qkrylov_operator_apply_device_f64(qkrylov_operator_h op, const void* x_dev, void* y_dev, int nvec,
                                      uint64_t ld_in, uint64_t ld_out, void* stream);

/* Full complex diagonal, accumulated per term in OpSum order (reference
 * cpp:81-89); diag_complex: 2*dim doubles. */
int qkrylov_hamiltonian_diagonal_f64(qkrylov_hamiltonian_h h, double* diag_complex);

/* Compiled program and kernel choice as JSON (groups per kind, tiles, grid,
 * exchange path, states per lane, bond batch, block grouping). */
int qkrylov_hamiltonian_info(qkrylov_hamiltonian_h h, char* buf, size_t len);

/* Autotuner: time the apply-kernel variants of h (states per lane, bond
 * batch, and for nvec > 1 the vectors per block launch) on scratch vectors
 * and keep the fastest for later applies; buf receives the timings (JSON,
 * may be NULL).  Allocates 2*nvec*dim complex doubles while it runs. */
int qkrylov_hamiltonian_autotune(qkrylov_hamiltonian_h h, int nvec, char* buf, size_t len);

/* Lanczos ground state in FP64 (reference src/solvers/lanczos.cpp:101-202).
 * start_complex: NULL for the reference start vector (std::mt19937(1234),
 * uniform(-1,1), imaginary part drawn before the real part per element, as
 * the reference's GCC build evaluates KComplex(dist(rng), dist(rng))), or a
 * caller vector of 2*dim doubles.  eigvec_complex may be NULL. */
int qkrylov_lanczos_ground_state_f64(qkrylov_hamiltonian_h h, int maxiter, double tol,
                                     const double* start_complex,
                                     qkrylov_lanczos_result_f64_t* result,
                                     double* eigvec_complex);

/* As qkrylov_lanczos_ground_state_f64, with checkpoints of the run (no
 * reference counterpart: the reference keeps its basis in memory only,
 * src/solvers/lanczos.cpp:123-126, 169-171).  save_path (may be NULL):
 * file that receives the run's state (every stored basis vector, alpha,
 * beta, the last energy; raw binary, checksummed, replaced atomically)
 * every save_every steps (0: never) and when the run stops at stop_after;
 * stop_after > 0 ends the run once that many steps are done (result
 * converged = 0, iterations = stop_after), as a preempted job would;
 * resume_path (may be NULL): continue the run saved there instead of
 * starting one (start is then ignored).  A resumed run performs the same
 * arithmetic as an uninterrupted one: the same energy and iteration count.
 * The checkpoint records the Hamiltonian's fingerprint and dimension; a
 * mismatch is an error.  Host memory of (steps + 1) * dim complex doubles is
 * used while loading. */
int qkrylov_lanczos_checkpointed_f64(qkrylov_hamiltonian_h h, int maxiter, double tol,
                                     const double* start_complex, const char* save_path,
                                     int save_every, int stop_after, const char* resume_path,
                                     qkrylov_lanczos_result_f64_t* result,
                                     double* eigvec_complex);

/* Eigenpair files (what the reference's examples/hdf5_storage.py:5-13 keeps
 * in HDF5, which this build does not depend on): n eigenvalues and, when
 * vectors_complex is not NULL, n vectors of dim complex doubles one after
 * another; fingerprint: qkrylov_hamiltonian_fingerprint() of the
 * Hamiltonian they belong to, or 0.  Raw binary, checksummed. */
int qkrylov_eigenpairs_save_f64(const char* path, uint64_t dim, int n, const double* values,
                                const double* vectors_complex, uint64_t fingerprint);
/* Header of an eigenpair file (any output pointer may be NULL). */
int qkrylov_eigenpairs_info(const char* path, uint64_t* dim, int* n, int* has_vectors,
                            uint64_t* fingerprint);
/* values_out: n doubles; vectors_complex_out: n * 2 * dim doubles or NULL. */
int qkrylov_eigenpairs_load_f64(const char* path, double* values_out, double* vectors_complex_out);
/* Hash of what defines h (site kind, size, sector, dimension, every OpSum
 * term in order); 0 on error. */
uint64_t qkrylov_hamiltonian_fingerprint(qkrylov_hamiltonian_h h);

/* Block Lanczos with full reorthogonalisation (no reference counterpart; it
 * extends lanczos.cpp:101-202 to blocks of `block` <= 16 vectors): lowest nev
 * <= block Ritz values, converged when each moved by less than tol in one
 * block step.  evecs_complex_out: nev*2*dim doubles or NULL. */
int qkrylov_block_lanczos_f64(qkrylov_hamiltonian_h h, int block, int nev, int maxiter, double tol,
                              double* eigenvalues_out, double* eigenvectors_complex_out,
                              int* iterations_out, int* converged_out);

/* Davidson lowest eigenpairs in FP64 (reference src/solvers/davidson.cpp:85-227),
 * Jacobi diagonal correction, at most 100 iterations as the reference.
 * eigenvalues_out[n_eig]; eigenvectors_complex_out: n_eig*2*dim doubles or NULL. */
int qkrylov_davidson_lowest_f64(qkrylov_hamiltonian_h h, int n_eig, int max_subspace, double tol,
                                double* eigenvalues_out, double* eigenvectors_complex_out,
                                qkrylov_davidson_result_c_t* result_info);

/* ---- Learned Davidson preconditioner.  No counterpart in the reference,
 * whose Davidson corrects with the Jacobi diagonal (davidson.cpp:166-177).
 * With J_i = 1/(H_ii - theta) (clamped as the reference does), z = J r,
 * w = J (H z - diag(H) z) and w2 = J (H w - diag(H) w), the correction is
 * t_i = z_i + s1_i w_i + s2_i w2_i with s1 = -0.5 + 0.5 tanh(a),
 * s2 = 0.5 tanh(b), where (a_i, b_i) is the output of a 3-layer MLP
 * (32 -> 128 -> 128 -> 2, ReLU, BF16 weights, FP32 accumulation on MFMA) on
 * per-state features (diagonal, theta, connectivity, local occupations).
 * (s1, s2) = (-1, 1) would be the second-order Neumann expansion of
 * (H - theta)^-1 r; an untrained network (last layer 0) gives the damped
 * first-order correction t = z - w/2.  Each correction costs two extra
 * applies of H. ---- */
typedef struct qkrylov_mlp_t* qkrylov_mlp_h;

typedef struct {
    int pairs;             /* (residual, accurate correction) pairs generated */
    int epochs;
    double loss_first;     /* mean loss (qkrylov_mlp_pair_loss_f64) before training */
    double loss_last;      /* after training */
    double solve_rel_res;  /* worst relative residual of MINRES-solved corrections
                              (0: the exact image K t* = -r was used) */
} qkrylov_mlp_train_info_t;

/* Random hidden layers from std::mt19937(seed), output layer zero. */
qkrylov_mlp_h qkrylov_mlp_create(unsigned seed);
void qkrylov_mlp_destroy(qkrylov_mlp_h m);
/* Parameter count (20994) and FP32 layout: W1[32][128], b1[128],
 * W2[128][128], b2[128], W3[128][2], b3[2] (row index = input). */
int qkrylov_mlp_num_params(void);
int qkrylov_mlp_get_params(qkrylov_mlp_h m, float* params_out);
int qkrylov_mlp_set_params(qkrylov_mlp_h m, const float* params);
/* o = net(features) (n x 2: a, b per row) for n rows of 32 features
 * (rounded to BF16 on entry); with grad_out (n x 2) and grad_params
 * non-NULL also grad_params = d(sum_ik grad_out[i][k] o[i][k]) / d(params).
 * Host buffers. */
int qkrylov_mlp_eval(qkrylov_mlp_h m, const float* features, uint64_t n, const float* grad_out,
                     float* o_out, float* grad_params);
/* Features of every basis state of h for theta: dim x 32 floats (the BF16
 * values the network reads). */
int qkrylov_mlp_features(qkrylov_hamiltonian_h h, double theta, float* features_out);
/* The learned correction t of a host residual r (complex, dim) at theta, as
 * Davidson applies it; z_out / w_out / w2_out (may be NULL) receive the
 * three terms. */
int qkrylov_mlp_correction_f64(qkrylov_mlp_h m, qkrylov_hamiltonian_h h, const double* r_complex,
                               double theta, double* t_complex_out, double* z_complex_out,
                               double* w_complex_out, double* w2_complex_out);
/* Training loss of npairs pairs on h's basis (host buffers, pair-major, dim
 * states each): o (npairs*dim x 2 floats, network output), z, w, w2, b, u
 * (complex), theta (npairs doubles).  With K = P_u (H - theta) P_u,
 * P_u = 1 - u u^H (u unit) and t = P_u (z + s1 w + s2 w2):
 * loss = sum over pairs of 1 - |<b, K t>|^2 / |K t|^2, b = K t* / |K t*|
 * for the accurate correction t*; grad_o (npairs*dim x 2, may be NULL) =
 * d(loss / npairs) / d(o). */
int qkrylov_mlp_pair_loss_f64(qkrylov_hamiltonian_h h, int npairs, const float* o,
                              const double* z_complex, const double* w_complex,
                              const double* w2_complex, const double* b_complex,
                              const double* u_complex, const double* theta, float* grad_o,
                              double* loss_out);
/* Accurate Jacobi-Davidson correction: MINRES on (I-uu^H)(H-theta)(I-uu^H) t
 * = -r, t orthogonal to u (host complex vectors of dim). */
int qkrylov_jd_correction_f64(qkrylov_hamiltonian_h h, const double* r_complex,
                              const double* u_complex, double theta, int max_iter, double rtol,
                              double* t_complex_out, double* rel_res_out);
/* Train m on pairs generated by a Jacobi Davidson run (n_eig, max_subspace,
 * max_iter, tol 1e-8) on h_small.  loss_history: epochs+1 doubles or NULL. */
int qkrylov_mlp_train_f64(qkrylov_mlp_h m, qkrylov_hamiltonian_h h_small, int n_eig,
                          int max_subspace, int max_iter, int max_pairs, int epochs,
                          int batch_pairs, double lr, double* loss_history,
                          qkrylov_mlp_train_info_t* info);
/* Davidson with an iteration limit and an optional learned preconditioner
 * (mlp NULL: the Jacobi diagonal, as qkrylov_davidson_lowest_f64). */
int qkrylov_davidson_lowest_ex_f64(qkrylov_hamiltonian_h h, int n_eig, int max_subspace,
                                   double tol, int max_iter, qkrylov_mlp_h mlp,
                                   double* eigenvalues_out, double* eigenvectors_complex_out,
                                   qkrylov_davidson_result_c_t* result_info);

/* Continued-fraction coefficients in FP64 (reference src/solvers/dynamics.cpp:12-59). */
int qkrylov_continued_fraction_coeffs_f64(qkrylov_hamiltonian_h h, const double* phi0_complex,
                                          int n_iter, double* alphas_out, double* betas_out,
                                          double* norm_phi0_out, int* num_coeffs_out);

/* nblock independent continued-fraction recurrences run as one block: phi0
 * vector b at offset b*2*dim doubles; alphas_out/betas_out: nblock*n_iter;
 * each row equals what qkrylov_continued_fraction_coeffs_f64 returns for that
 * start vector. */
int qkrylov_continued_fraction_block_f64(qkrylov_hamiltonian_h h, const double* phi0_complex,
                                         int nblock, int n_iter, double* alphas_out,
                                         double* betas_out, double* norms_out, int* counts_out);

/* As qkrylov_continued_fraction_block_f64 with the start vectors already in
 * device memory: vector b at phi0_dev + b*dim complex doubles. */
int qkrylov_continued_fraction_block_device_f64(qkrylov_hamiltonian_h h, const void* phi0_dev,
                                                int nblock, int n_iter, double* alphas_out,
                                                double* betas_out, double* norms_out,
                                                int* counts_out);

/* Reference src/solvers/dynamics.cpp:61-89 in FP64. */
double qkrylov_evaluate_spectral_function_f64(const double* alphas, const double* betas, size_t n,
                                              double norm_phi0, double omega, double E0, double eta);

/* Host-only: eigen-decomposition of a small Hermitian matrix (row-major
 * n x n complex), ascending eigenvalues, vecs_out[k*n + j] = component j of
 * eigenvector k; the solver of the Davidson projected matrix (reference
 * davidson.cpp:22-82 uses Jacobi rotations; any converged solver gives the
 * same eigenvalues).  Householder tridiagonalisation and implicit-shift QL. */
int qkrylov_hermitian_eigh_f64(int n, const double* a_complex, double* evals_out,
                               double* vecs_complex_out);

/* FTLM for one beta in FP64 (reference src/solvers/ftlm.cpp:122-170):
 * std::mt19937(42) re-seeded per call, normal(0,1), imaginary part drawn
 * before the real part per element. */
int qkrylov_ftlm_f64(qkrylov_hamiltonian_h h, double beta, int n_random, int n_steps,
                     qkrylov_ftlm_result_f64_t* result);

/* FTLM over several sectors and a temperature grid.  For sector s and beta b
 * the outputs Z[s*nbeta+b], E[...], Cv[...] equal qkrylov_ftlm_f64(hs[s],
 * betas[b], n_random, n_steps).  mode 0: every (sector, random vector) pair
 * advanced together in one block recurrence per step; modes 1-3: one vector
 * at a time (the plain loop).  mode 1: each vector through mode 0's apply
 * kernel, alpha from its partial sums per tile (the same arithmetic per
 * vector as mode 0 for sectors below 2^22 states); mode 2: alpha from a
 * separate dot as the reference's loop takes it (ftlm.cpp:24-53); mode 3:
 * the single-vector apply kernel, alpha from its partial sums per workgroup.
 * Other modes: QKRYLOV_ERROR_INVALID_ARG. */
int qkrylov_ftlm_sweep_f64(const qkrylov_hamiltonian_h* hs, int nsec, const double* betas, int nbeta,
                           int n_random, int n_steps, int mode,
                           double* Z_out, double* E_out, double* Cv_out);

/* phi_q = N^-1/2 sum_j exp(i q_k j) Sz_j psi for q_k = 2*pi*k/nsites,
 * k = 0..nq-1, on device buffers: psi (dim complex), phi (nq*ld complex).
 * Spin-1/2 bases only. */
int qkrylov_sz_q_device_f64(qkrylov_hamiltonian_h h, const void* psi_dev, void* phi_dev,
                            int nq, uint64_t ld, void* stream);

/* G[i*nw + j] = sum_k conj(V_i[k]) W_j[k] on the GPU (FP64 MFMA), the
 * tall-skinny projection Davidson uses for new rows of V^H H V: host
 * vectors V (m*2*n doubles, m <= 64) and W (nw*2*n doubles, nw <= 16). */
int qkrylov_block_dot_f64(const double* V_complex, int m, const double* W_complex, int nw,
                          uint64_t n, double* G_complex_out);

/* G = V^H W as qkrylov_block_dot_f64 computes it (same kernels, bit for
 * bit), with device buffers and no allocation of any kind: no reference
 * counterpart; for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * V_i = V_dev + i*ldv and W_j = W_dev + j*ldw complex doubles (i < m <= 64,
 * j < nw <= 16, ldv, ldw >= n); G_dev receives m*nw complex doubles,
 * G[i*nw + j] = sum_k conj(V_i[k]) W_j[k].  Asynchronous on stream (a
 * hipStream_t, NULL: the default stream).  Its scratch (pointer tables and
 * partial sums) is one device array of 16 MiB + 1 KiB per GPU, held until
 * the process ends and shared by every call on that GPU: calls that overlap
 * in time on one GPU race.  Sizes beyond the array (none within the m, nw
 * limits) return QKRYLOV_ERROR_INVALID_ARG. */
int qkrylov_block_dot_noalloc_f64(const void* V_dev, uint64_t ldv, int m, const void* W_dev,
                                  uint64_t ldw, int nw, uint64_t n, void* G_dev, void* stream);

/* Synchronise the handle's stream. */
int qkrylov_hamiltonian_sync(qkrylov_hamiltonian_h h);

/* The reference solvers' random streams, for callers that want the same
 * start vectors: out receives n complex values (2n doubles).
 * start: std::mt19937(seed), uniform(-1,1) (lanczos.cpp:115-120);
 * normals: std::mt19937(seed), normal(0,1), one stream (ftlm.cpp:133-143).
 * In both, the imaginary part of an element is drawn before its real part. */
int qkrylov_reference_start_f64(uint64_t n, unsigned seed, double* out);
int qkrylov_reference_normals_f64(uint64_t n, unsigned seed, double* out);

#ifdef __cplusplus
}
#endif

#endif /* QKRYLOV_C_API_H */
