// Krylov solvers on one GPU (FP64 vectors in HBM, device-resident scalars,
// host small eigensolves).  Semantics follow the reference solvers
// (src/solvers/lanczos.cpp, davidson.cpp, ftlm.cpp, dynamics.cpp).
#pragma once

#include "qk_hamiltonian.hpp"

#include <string>
#include <vector>

namespace qk {

struct LanczosResult {
    double energy = 0.0;
    int iterations = 0;
    bool converged = false;
    bool stopped = false;            // ended at LanczosIO::stop_after (not converged)
    int second_passes = 0;           // steps that ran the second Gram-Schmidt pass
    std::vector<cplx> eigenvector;   // host, empty unless requested
};

// Checkpointing of a Lanczos run.  save_path: write the run's state (every
// stored basis vector, alpha, beta, the last energy) every save_every steps
// and at stop_after; stop_after > 0: end the run once that many steps are
// done (as a preempted job would), result.stopped = true; resume_path:
// continue a run from its checkpoint instead of starting it.  A resumed run
// performs the same arithmetic as an uninterrupted one (same energy, same
// iteration count).
struct LanczosIO {
    std::string save_path;
    int save_every = 0;
    int stop_after = 0;
    std::string resume_path;
};

// start: host vector of dim() entries, or null for the reference start
// (std::mt19937(1234), uniform(-1,1)).
LanczosResult lanczos_ground_state(const Hamiltonian& H, int maxiter, double tol,
                                   const cplx* start, bool want_vector);
LanczosResult lanczos_run(const Hamiltonian& H, int maxiter, double tol, const cplx* start,
                          bool want_vector, const LanczosIO& io);

struct BlockLanczosResult {
    std::vector<double> eigenvalues;               // lowest nev Ritz values
    std::vector<std::vector<cplx>> eigenvectors;   // host, empty unless requested
    int iterations = 0;                            // block steps
    bool converged = false;
};

// Block Lanczos with full reorthogonalisation: blocks of b vectors (b <= 16),
// H applied to the whole block per step, block three-term recurrence
// H Q_j = Q_{j-1} B_j^H + Q_j A_j + Q_{j+1} B_{j+1}, projections on FP64 MFMA,
// Cholesky-QR twice for each new block.  Converged when every one of the
// lowest nev Ritz values moved by less than tol in one step (the
// reference's Lanczos rule, lanczos.cpp:173-182, per value).  The first
// start vector is the reference's (std::mt19937(1234), uniform(-1,1)); the
// others continue that stream.
BlockLanczosResult block_lanczos_lowest(const Hamiltonian& H, int b, int nev, int maxiter,
                                        double tol, bool want_vectors);

struct DavidsonResult {
    std::vector<double> eigenvalues;
    std::vector<std::vector<cplx>> eigenvectors;   // host, empty unless requested
    int iterations = 0;
    bool converged = false;
    std::vector<double> residuals;                 // final residual norms
};

// Preconditioner hook: t = M(r; lambda).  When null, Jacobi (reference).
struct Preconditioner {
    virtual ~Preconditioner() = default;
    // r, t: device vectors of dim entries; diag: device complex diagonal;
    // u: the Ritz vector of lambda when wants_ritz(), else null.
    virtual void apply(const Hamiltonian& H, const cplx* r, const cplx* diag, double lambda,
                       const cplx* u, cplx* t, hipStream_t st) = 0;
    // The nt corrections of one Davidson iteration (host arrays of device
    // pointers / values); the default applies them one by one.
    virtual void apply_many(const Hamiltonian& H, int nt, const cplx* const* r, const cplx* diag,
                            const double* lambda, const cplx* const* u, cplx* const* t,
                            hipStream_t st) {
        for (int j = 0; j < nt; ++j) apply(H, r[j], diag, lambda[j], u ? u[j] : nullptr, t[j], st);
    }
    virtual bool wants_ritz() const { return false; }
};

// max_iter: the reference stops after 100 iterations (davidson.cpp:141).
DavidsonResult davidson_lowest(const Hamiltonian& H, int n_eig, int max_subspace, double tol,
                               bool want_vectors, Preconditioner* pre = nullptr,
                               int max_iter = 100);

struct CFResult {
    std::vector<double> alphas, betas;
    double norm_phi0 = 0.0;
};

// nblock start vectors (host, vector b at offset b*dim) as one block.
std::vector<CFResult> continued_fraction_block(const Hamiltonian& H, const cplx* phi0, int nblock,
                                               int n_iter);
// Same, with the start vectors already on the device (ld = dim).
std::vector<CFResult> continued_fraction_block_dev(const Hamiltonian& H, const cplx* phi0_dev,
                                                   int nblock, int n_iter);

double spectral_function(const double* alphas, const double* betas, size_t n, double norm_phi0,
                         double omega, double E0, double eta);

struct FTLMPoint {
    double beta, Z, E, Cv;
};

// ftlm over sectors and a beta grid; out[s*nbeta + b].  mode 0: all
// (sector, vector) pairs advanced together; modes 1-3 one vector at a time:
// 1: through mode 0's multi-sector kernel, alpha from its partial sums per
// tile (the same arithmetic per vector as mode 0 where the sectors take that
// launch); 2: alpha from a separate pass over the vector and its image, as
// the reference's loop takes it (ftlm.cpp:24-53); 3: the single-vector
// kernel, alpha from its partial sums per workgroup.
std::vector<FTLMPoint> ftlm_sweep(const std::vector<const Hamiltonian*>& hs,
                                  const std::vector<double>& betas, int n_random, int n_steps,
                                  int mode);

}  // namespace qk
