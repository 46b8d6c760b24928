// Learned Davidson preconditioner: training-pair generation, accurate
// corrections, the training loop and the preconditioner used by Davidson.
//
// Pairs come from the solver itself: a Jacobi Davidson run on a small
// lattice records, for every correction it computes, the residual r, the
// Ritz value theta and the Ritz vector u.  The accurate correction t* solves
// the Jacobi-Davidson correction equation
//     (I - u u^H)(H - theta)(I - u u^H) t = -r,   t orthogonal to u,
// by MINRES (Paige & Saunders 1975) on the GPU; it is stored normalised.
#include "qk_learned.hpp"

#include "qk_vec.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>

namespace qk {

namespace {

template <typename T>
T fetch1(const void* dev, hipStream_t st) {
    T v;
    QK_HIP(hipMemcpyAsync(&v, dev, sizeof(T), hipMemcpyDeviceToHost, st));
    QK_HIP(hipStreamSynchronize(st));
    return v;
}

struct Blas {
    hipStream_t st;
    u64 n;
    DevBuf red, sc{64 * sizeof(double)};
    cplx dot(const cplx* a, const cplx* b) {   // sum conj(a) b
        vec::dot(a, b, n, sc.get<cplx>(), red, st);
        return fetch1<cplx>(sc.get(), st);
    }
    double norm(const cplx* a) {
        vec::norm2(a, n, sc.get<double>(), red, st);
        return std::sqrt(fetch1<double>(sc.get(), st));
    }
    void axpy(cplx a, const cplx* x, cplx* y) { vec::axpy(a, x, y, n, st); }
    void scale(cplx* x, double s) { vec::scale(x, n, cplx{s, 0.0}, st); }
    void copy(cplx* y, const cplx* x) {
        QK_HIP(hipMemcpyAsync(y, x, sizeof(cplx) * n, hipMemcpyDeviceToDevice, st));
    }
    void zero(cplx* x) { QK_HIP(hipMemsetAsync(x, 0, sizeof(cplx) * n, st)); }
};

// y = (I - u u^H)(H - theta) x for x orthogonal to u
void proj_op(const Hamiltonian& H, Blas& b, const cplx* u, double theta, const cplx* x, cplx* y) {
    H.apply(x, y, 1, b.n, b.st);
    b.axpy(cplx{-theta, 0.0}, x, y);
    const cplx c = b.dot(u, y);
    b.axpy(cplx{-c.re, -c.im}, u, y);
}

}  // namespace

// ---------------------------------------------------------------------------
// MINRES for the projected correction equation; returns the relative residual.
// ---------------------------------------------------------------------------
double jd_correction(const Hamiltonian& H, const cplx* r, const cplx* u, double theta, cplx* t,
                     int max_it, double rtol, hipStream_t st) {
    const u64 n = H.dim();
    Blas b{st, n};
    DevBuf buf(sizeof(cplx) * n * 6);
    cplx* v_prev = buf.get<cplx>();
    cplx* v = v_prev + n;
    cplx* p = v + n;
    cplx* w_prev2 = p + n;
    cplx* w_prev = w_prev2 + n;
    cplx* w = w_prev + n;
    b.zero(t);
    b.zero(v_prev);
    b.zero(w_prev2);
    b.zero(w_prev);
    // right-hand side -r (orthogonal to u for a Ritz pair; projected anyway)
    b.copy(v, r);
    b.scale(v, -1.0);
    {
        const cplx c = b.dot(u, v);
        b.axpy(cplx{-c.re, -c.im}, u, v);
    }
    const double beta1 = b.norm(v);
    if (!(beta1 > 0.0)) return 0.0;
    b.scale(v, 1.0 / beta1);
    double beta = 0.0, eta = beta1;
    double c_prev = 1.0, s_prev = 0.0, c_prev2 = 1.0, s_prev2 = 0.0;
    double res = 1.0;
    for (int k = 0; k < max_it; ++k) {
        proj_op(H, b, u, theta, v, p);
        const double alpha = b.dot(v, p).re;
        b.axpy(cplx{-alpha, 0.0}, v, p);
        if (k > 0) b.axpy(cplx{-beta, 0.0}, v_prev, p);
        const double beta_next = b.norm(p);
        // previous rotations applied to the new column (beta, alpha, beta_next)
        const double eps = s_prev2 * beta;
        const double dbar = c_prev2 * beta;
        const double delta = c_prev * dbar + s_prev * alpha;
        const double gbar = -s_prev * dbar + c_prev * alpha;
        const double gamma = std::sqrt(gbar * gbar + beta_next * beta_next);
        if (!(gamma > 0.0)) break;
        const double cc = gbar / gamma, ss = beta_next / gamma;
        // w = (v - delta w_prev - eps w_prev2) / gamma
        b.copy(w, v);
        b.axpy(cplx{-delta, 0.0}, w_prev, w);
        b.axpy(cplx{-eps, 0.0}, w_prev2, w);
        b.scale(w, 1.0 / gamma);
        b.axpy(cplx{cc * eta, 0.0}, w, t);
        eta = -ss * eta;
        res = std::fabs(eta) / beta1;
        // shift
        std::swap(w_prev2, w_prev);
        std::swap(w_prev, w);
        c_prev2 = c_prev;
        s_prev2 = s_prev;
        c_prev = cc;
        s_prev = ss;
        if (res < rtol || !(beta_next > 1e-300)) break;
        std::swap(v_prev, v);
        std::swap(v, p);   // v = p / beta_next
        b.scale(v, 1.0 / beta_next);
        beta = beta_next;
    }
    return res;
}

// ---------------------------------------------------------------------------
// Jacobi preconditioner that records (r, theta, u) of every correction
// ---------------------------------------------------------------------------
namespace {

struct Recorder : Preconditioner {
    int max_pairs;
    u64 n;
    std::vector<DevBuf> r, u;
    std::vector<double> theta;
    Recorder(int mp, u64 dim) : max_pairs(mp), n(dim) {}
    bool wants_ritz() const override { return true; }
    void apply(const Hamiltonian&, const cplx* rr, const cplx* diag, double lambda, const cplx* uu,
               cplx* t, hipStream_t st) override {
        vec::jacobi_precond(rr, diag, lambda, t, n, st);
        if (int(theta.size()) >= max_pairs || !uu) return;
        r.emplace_back(sizeof(cplx) * n);
        u.emplace_back(sizeof(cplx) * n);
        QK_HIP(hipMemcpyAsync(r.back().get(), rr, sizeof(cplx) * n, hipMemcpyDeviceToDevice, st));
        QK_HIP(hipMemcpyAsync(u.back().get(), uu, sizeof(cplx) * n, hipMemcpyDeviceToDevice, st));
        theta.push_back(lambda);
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// training
// ---------------------------------------------------------------------------
MlpTrainStats train_mlp(Mlp& net, const Hamiltonian& Hs, const MlpTrainConfig& cfg,
                        std::vector<double>* loss_history) {
    MlpTrainStats out;
    const u64 n = Hs.dim();
    if (n == 0) return out;
    hipStream_t st = Hs.stream;

    // 1. pairs from a Jacobi Davidson run on the small lattice
    Recorder rec(cfg.max_pairs, n);
    (void)davidson_lowest(Hs, cfg.n_eig, cfg.max_subspace, cfg.tol, false, &rec, cfg.max_iter);
    const int P = int(rec.theta.size());
    out.pairs = P;
    if (P == 0) return out;

    // 2. accurate corrections t*, stored as the unit vector K t* (the loss
    // target, K = P_u (H - theta) P_u): K t* = -r for the exact solution, or
    // K applied to a MINRES solution when cfg.solve_iters > 0; and the three
    // correction terms z = J r, w = J (H z - diag z), w2 = J (H w - diag w)
    // of every pair; everything pair-major
    const std::shared_ptr<const MlpStates> states_p = mlp_states_of(Hs);
    const MlpStates& states = *states_p;
    DevBuf Z(sizeof(cplx) * n * P), W(sizeof(cplx) * n * P), W2(sizeof(cplx) * n * P),
        T(sizeof(cplx) * n * P), U(sizeof(cplx) * n * P), hz(sizeof(cplx) * n), ts(sizeof(cplx) * n);
    Blas b{st, n};
    double worst = 0.0;
    for (int p = 0; p < P; ++p) {
        cplx* tp = T.get<cplx>() + u64(p) * n;
        cplx* zp = Z.get<cplx>() + u64(p) * n;
        const cplx* up = rec.u[p].get<cplx>();
        if (cfg.solve_iters > 0) {
            const double rr = jd_correction(Hs, rec.r[p].get<cplx>(), up, rec.theta[p], ts.get<cplx>(),
                                            cfg.solve_iters, cfg.solve_rtol, st);
            worst = std::max(worst, rr);
            proj_op(Hs, b, up, rec.theta[p], ts.get<cplx>(), tp);
        } else {
            // K t* = -r exactly for the solution t* of the correction equation
            b.copy(tp, rec.r[p].get<cplx>());
            b.scale(tp, -1.0);
            const cplx c = b.dot(up, tp);
            b.axpy(cplx{-c.re, -c.im}, up, tp);
        }
        const double nt = b.norm(tp);
        if (nt > 0.0) b.scale(tp, 1.0 / nt);
        cplx* wp = W.get<cplx>() + u64(p) * n;
        mlp_jacobi(states, rec.theta[p], rec.r[p].get<cplx>(), zp, st);
        Hs.apply(zp, hz.get<cplx>(), 1, n, st);
        mlp_neumann(states, rec.theta[p], hz.get<cplx>(), zp, wp, st);
        Hs.apply(wp, hz.get<cplx>(), 1, n, st);
        mlp_neumann(states, rec.theta[p], hz.get<cplx>(), wp, W2.get<cplx>() + u64(p) * n, st);
        QK_HIP(hipMemcpyAsync(U.get<cplx>() + u64(p) * n, rec.u[p].get(), sizeof(cplx) * n,
                              hipMemcpyDeviceToDevice, st));
    }
    QK_HIP(hipStreamSynchronize(st));
    rec.r.clear();
    rec.u.clear();
    out.solve_rel_res = worst;

    // 3. features, then Adam over mini-batches of consecutive pairs
    const int bp = std::max(1, std::min(cfg.batch_pairs, P));
    DevBuf X(sizeof(u16) * n * P * kMlpIn), o(sizeof(float) * n * bp * kMlpOut),
        go(sizeof(float) * n * bp * kMlpOut), h1(sizeof(u16) * n * bp * kMlpHid),
        h2(sizeof(u16) * n * bp * kMlpHid), grads(sizeof(float) * kMlpParams);
    MlpLossWork wk;
    for (int p = 0; p < P; ++p) mlp_features(states, rec.theta[p], X.get<u16>() + u64(p) * n * kMlpIn, st);
    DevBuf TH(sizeof(double) * P);   // the pairs' Ritz values, for the loss
    QK_HIP(hipMemcpyAsync(TH.get(), rec.theta.data(), sizeof(double) * P, hipMemcpyHostToDevice, st));
    const double* th = TH.get<double>();

    auto eval_all = [&]() {
        double L = 0.0;
        for (int p0 = 0; p0 < P; p0 += bp) {
            const int cnt = std::min(bp, P - p0);
            net.forward(X.get<u16>() + u64(p0) * n * kMlpIn, n * cnt, o.get<float>(), nullptr, nullptr, st);
            L += mlp_pair_loss(Hs, Z.get<cplx>() + u64(p0) * n, W.get<cplx>() + u64(p0) * n,
                               W2.get<cplx>() + u64(p0) * n, T.get<cplx>() + u64(p0) * n, U.get<cplx>() + u64(p0) * n,
                               th + p0, cnt, o.get<float>(), nullptr, wk, st);
        }
        return L / P;
    };
    out.loss_first = eval_all();
    if (loss_history) loss_history->push_back(out.loss_first);
    std::mt19937 rng(cfg.seed);
    std::vector<int> starts;
    for (int p0 = 0; p0 < P; p0 += bp) starts.push_back(p0);
    for (int e = 0; e < cfg.epochs; ++e) {
        std::shuffle(starts.begin(), starts.end(), rng);
        for (int p0 : starts) {
            const int cnt = std::min(bp, P - p0);
            const u16* Xb = X.get<u16>() + u64(p0) * n * kMlpIn;
            net.forward(Xb, n * cnt, o.get<float>(), h1.get<u16>(), h2.get<u16>(), st);
            (void)mlp_pair_loss(Hs, Z.get<cplx>() + u64(p0) * n, W.get<cplx>() + u64(p0) * n,
                                W2.get<cplx>() + u64(p0) * n, T.get<cplx>() + u64(p0) * n, U.get<cplx>() + u64(p0) * n,
                                th + p0, cnt, o.get<float>(), go.get<float>(), wk, st, false);
            net.backward(Xb, h1.get<u16>(), h2.get<u16>(), go.get<float>(), n * cnt,
                         grads.get<float>(), st);
            net.adam_step(grads.get<float>(), float(cfg.lr), st);
        }
        out.epochs = e + 1;
        if (loss_history) loss_history->push_back(eval_all());
    }
    out.loss_last = eval_all();
    QK_HIP(hipStreamSynchronize(st));
    return out;
}

// ---------------------------------------------------------------------------
// the preconditioner
// ---------------------------------------------------------------------------
MlpPreconditioner::MlpPreconditioner(const Mlp& net, const Hamiltonian& H)
    : net_(net), sp_(mlp_states_of(H)), states_(*sp_) {}

void MlpPreconditioner::apply(const Hamiltonian& H, const cplx* r, const cplx* diag, double lambda,
                              const cplx* u, cplx* t, hipStream_t st) {
    cplx* tp[1] = {t};
    const cplx* up[1] = {u};
    apply_many(H, 1, &r, diag, &lambda, u ? up : nullptr, tp, st);
}

void MlpPreconditioner::apply_many(const Hamiltonian& H, int nt, const cplx* const* r, const cplx*,
                                   const double* lambda, const cplx* const* u, cplx* const* t,
                                   hipStream_t st) {
    const u64 n = states_.n;
    if (!n || nt <= 0) return;
    const u64 tot = n * u64(nt);
    // network rows per correction: one per class of equal inputs, or one per state
    const u64 rows = states_.ncls ? states_.ncls : n;
    X_.resize(sizeof(u16) * rows * nt * kMlpIn);
    o_.resize(sizeof(float) * rows * nt * kMlpOut);
    z_.resize(sizeof(cplx) * tot);
    w_.resize(sizeof(cplx) * tot);
    w2_.resize(sizeof(cplx) * tot);
    cplx* z = z_.get<cplx>();
    cplx* w = w_.get<cplx>();
    cplx* w2 = w2_.get<cplx>();
    // the corrections in batches of up to kMlpBatch, one launch per stage
    for (int j0 = 0; j0 < nt; j0 += kMlpBatch) {
        const int nb = std::min(kMlpBatch, nt - j0);
        MlpBatch b{};
        b.nt = nb;
        for (int j = 0; j < nb; ++j) {
            b.r[j] = r[j0 + j];
            b.t[j] = t[j0 + j];
            b.u[j] = u ? u[j0 + j] : nullptr;
            b.theta[j] = lambda[j0 + j];
        }
        cplx* zb = z + u64(j0) * n;
        cplx* wb = w + u64(j0) * n;
        cplx* w2b = w2 + u64(j0) * n;
        mlp_jacobi_many(states_, b, zb, st);
        H.apply(zb, wb, nb, n, st);   // w holds H z until the Neumann term overwrites it
        mlp_neumann_many(states_, b, wb, zb, wb, st);
        H.apply(wb, w2b, nb, n, st);
        mlp_neumann_many(states_, b, w2b, wb, w2b, st);
        u16* Xb = X_.get<u16>() + u64(j0) * rows * kMlpIn;
        float* ob = o_.get<float>() + u64(j0) * rows * kMlpOut;
        mlp_features_many(states_, b, Xb, st);
        net_.forward(Xb, rows * u64(nb), ob, nullptr, nullptr, st);
        // t approximates (H - lambda)^-1 r, which is u itself for an exact
        // Ritz pair: the u component is removed here, before Davidson
        // orthogonalises t against its basis
        mlp_combine_project_many(states_, b, ob, zb, wb, w2b, red_, st);
    }
}

}  // namespace qk
