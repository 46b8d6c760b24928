// Training and use of the MLP preconditioner for Davidson (qk_mlp.hpp).
#pragma once

#include "qk_mlp.hpp"
#include "qk_solvers.hpp"

#include <vector>

namespace qk {

struct MlpTrainConfig {
    int n_eig = 8;            // Davidson run that generates the pairs
    int max_subspace = 48;
    double tol = 1e-8;
    int max_iter = 40;
    int max_pairs = 256;
    // MINRES iterations per accurate correction; 0: use K t* = -r, exact for
    // the solution t* of the correction equation (the loss reads t* only
    // through K t*)
    int solve_iters = 0;
    double solve_rtol = 1e-8;
    int epochs = 30;
    int batch_pairs = 8;
    double lr = 1e-3;
    unsigned seed = 7;
};

// MINRES solve of (I - uu^H)(H - theta)(I - uu^H) t = -r, t orthogonal to u
// (all device vectors of H.dim()); returns the final relative residual.
double jd_correction(const Hamiltonian& H, const cplx* r, const cplx* u, double theta, cplx* t,
                     int max_it, double rtol, hipStream_t st);

// Generate pairs on Hs with a Jacobi Davidson run, compute their accurate
// corrections and train `net`; loss_history gets the mean loss before
// training and after each epoch.
MlpTrainStats train_mlp(Mlp& net, const Hamiltonian& Hs, const MlpTrainConfig& cfg,
                        std::vector<double>* loss_history);

// t = z + s1 w + s2 w2 with z = J r, w = J (H z - diag z), w2 = J (H w -
// diag w) and (s1, s2) from the network (qk_mlp.hpp), then the component
// along the Ritz vector u removed; two extra applies of H per correction.
class MlpPreconditioner : public Preconditioner {
public:
    MlpPreconditioner(const Mlp& net, const Hamiltonian& H);
    void apply(const Hamiltonian& H, const cplx* r, const cplx* diag, double lambda, const cplx* u,
               cplx* t, hipStream_t st) override;
    // all corrections of an iteration with two batched applies of H and one
    // forward pass over every (correction, state) row
    void apply_many(const Hamiltonian& H, int nt, const cplx* const* r, const cplx* diag,
                    const double* lambda, const cplx* const* u, cplx* const* t,
                    hipStream_t st) override;
    bool wants_ritz() const override { return true; }
    // the three correction terms of the last apply (device; the first
    // correction's after apply_many)
    const cplx* z() const { return z_.get<cplx>(); }
    const cplx* w() const { return w_.get<cplx>(); }
    const cplx* w2() const { return w2_.get<cplx>(); }

private:
    const Mlp& net_;
    std::shared_ptr<const MlpStates> sp_;   // H's tables (cached on H)
    const MlpStates& states_;
    DevBuf X_, o_, z_, w_, w2_, red_;
};

}  // namespace qk
