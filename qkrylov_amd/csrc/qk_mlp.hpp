// Learned preconditioner for the Davidson correction equation (H - theta) t = -r.
//
// With J_i = 1/(H_ii - theta) (the reference's clamped Jacobi factor),
// O = H - diag(H), z = J r, w = J (O z) and w2 = J (O w) (two extra applies),
// the correction is
//     t_i = z_i + s1_i w_i + s2_i w2_i,
//     s1_i = -0.5 + 0.5 tanh(a_i),  s2_i = 0.5 tanh(b_i),
// where (a_i, b_i) is the output of a three-layer MLP (32 -> 128 -> 128 -> 2,
// ReLU) reading per-state features (diagonal element, theta, connectivity,
// local occupations).  (s1, s2) = (-1, 1) is the second-order Neumann
// expansion of (D + O - theta)^-1 r = J r - J O J r + J O J O J r - ...; an
// untrained network (zero output layer) gives (-1/2, 0), t = z - w/2, the
// first-order expansion damped by one half; the network moves each state's
// two coefficients within s1 in [-1, 0], s2 in [-1/2, 1/2].  The component
// along the Ritz vector u is removed from t.
// Weights are kept in FP32 (master copy, Adam moments) and cast to BF16 for
// the kernels; every matrix product runs on gfx950 MFMA (BF16 inputs, FP32
// accumulation).  Training minimises, per (residual, accurate correction)
// pair, 1 - cos^2 between t and the accurate correction t* in the norm of
// the correction operator K = P_u (H - theta) P_u (mlp_pair_loss): a plain
// cosine weighs every error component alike and prefers undamped Neumann
// terms, whose high-energy errors stall Davidson; K's norm weighs them by
// their energy.
#pragma once

#include <memory>

#include "qk_device.hpp"

#include <vector>

namespace qk {

struct Hamiltonian;

constexpr int kMlpIn = 32;     // input features (zero-padded)
constexpr int kMlpHid = 128;   // hidden width
constexpr int kMlpOut = 2;     // (a, b) per state

// FP32 parameter block, offsets in floats.  W1[k][n], W2[k][n] and W3[k][n]
// are stored input-major (row k = input feature / hidden unit of the
// previous layer).
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpIn * kMlpHid;
constexpr int kMlpW2 = kMlpB1 + kMlpHid;
constexpr int kMlpB2 = kMlpW2 + kMlpHid * kMlpHid;
constexpr int kMlpW3 = kMlpB2 + kMlpHid;
constexpr int kMlpB3 = kMlpW3 + kMlpHid * kMlpOut;
constexpr int kMlpParams = kMlpB3 + kMlpOut;

// Features written per state (the rest of the 32 inputs are zero).
constexpr int kMlpFeatures = 10;

// Round-to-nearest-even float -> bfloat16 bits (NaN -> quiet NaN).
inline u16 f32_to_bf16(float f) {
    u32 u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return u16(u >> 16);
}
inline float bf16_to_f32(u16 b) {
    const u32 u = u32(b) << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

struct MlpTrainStats {
    int pairs = 0;            // (residual, correction) pairs generated
    int epochs = 0;
    double loss_first = 0.0;  // mean loss over all pairs before training
    double loss_last = 0.0;   // after training
    double solve_rel_res = 0.0;   // worst relative residual of the accurate corrections
};

class Mlp {
public:
    explicit Mlp(unsigned seed);
    ~Mlp();
    Mlp(const Mlp&) = delete;
    Mlp& operator=(const Mlp&) = delete;

    void get_params(float* host) const;       // kMlpParams floats
    void set_params(const float* host);

    // o[m*2 + k] = net(X)_k for n rows of kMlpIn BF16 features (device).
    // With h1/h2 non-null the BF16 activations are kept for backward().
    void forward(const u16* X, u64 n, float* o, u16* h1, u16* h2, hipStream_t st) const;
    // grads (device, kMlpParams floats) = d(sum_m,k go[m*2+k] o[m*2+k]) /
    // d(params), from the activations of forward().
    void backward(const u16* X, const u16* h1, const u16* h2, const float* go, u64 n,
                  float* grads, hipStream_t st);
    void adam_step(const float* grads, float lr, hipStream_t st);

    // Host-buffer forward/backward on given features (gradient checks).
    void eval_host(const float* X, u64 n, const float* go, float* o, float* grads);

    hipStream_t stream() const { return st_; }

    DevBuf params;            // FP32 master
    DevBuf m1, m2;            // Adam moments
    DevBuf w1t, w2t, w2k, w3b;   // BF16: W1^T [n][k], W2^T [n][k], W2 [k][n], W3 [k][2]
    int adam_t = 0;

private:
    void cast();
    hipStream_t st_ = nullptr;
    DevBuf part_, wg_, sum_;   // backward partial sums, their chunk sums
};

// Per-state data of one Hamiltonian that the features need, on the device.
struct MlpStates {
    DevBuf diag;   // double: Re H_ii
    DevBuf stat;   // float[4 + 1]: occupation fractions (4), connectivity
    u64 n = 0;
    int nsites = 1;
    // States with equal (diag, stat) have equal features for every theta, so
    // inference evaluates the network once per class of such states:
    // cls[i] is state i's class, rdiag/rstat the class's inputs.  ncls = 0
    // when classes would not save at least three quarters of the rows.
    DevBuf cls, rdiag, rstat;
    u64 ncls = 0;
    void build(const Hamiltonian& H, hipStream_t st);
};
// H's MlpStates: depend only on H (not on a network or theta), so they are
// built once per Hamiltonian and shared by every preconditioner and
// training run on it.
std::shared_ptr<const MlpStates> mlp_states_of(const Hamiltonian& H);

// The corrections of one Davidson iteration, passed by value to the batched
// kernels below (grid.y = correction): residuals r, outputs t, Ritz vectors
// u (null: none), shifts theta.
constexpr int kMlpBatch = 16;
struct MlpBatch {
    const cplx* r[kMlpBatch];
    cplx* t[kMlpBatch];
    const cplx* u[kMlpBatch];
    double theta[kMlpBatch];
    int nt;
};
// z_j = J_j r_j and, after hz_j = H z_j, w_j = J_j (hz_j - diag z_j) (w may
// be hz; the same call on (H w, w) gives w2); z, w hold the nt vectors
// contiguously (n each)
void mlp_jacobi_many(const MlpStates& s, const MlpBatch& b, cplx* z, hipStream_t st);
void mlp_neumann_many(const MlpStates& s, const MlpBatch& b, const cplx* hz, const cplx* z, cplx* w,
                      hipStream_t st);
// features of every network row of every correction: rows = ncls (classes)
// or n (states) per correction, correction j at X + j * rows * kIn
void mlp_features_many(const MlpStates& s, const MlpBatch& b, u16* X, hipStream_t st);
// t_j = z_j + s1 w_j + s2 w2_j from output rows j * rows + (class of the
// state or the state), then t_j -= <u_j, t_j> u_j where u_j is given
void mlp_combine_project_many(const MlpStates& s, const MlpBatch& b, const float* o, const cplx* z,
                              const cplx* w, const cplx* w2, DevBuf& scratch, hipStream_t st);

// X[i] = features of state i for theta (device).
void mlp_features(const MlpStates& s, double theta, u16* X, hipStream_t st);
// z = J r (device; J the reference's clamped Jacobi factor for theta)
void mlp_jacobi(const MlpStates& s, double theta, const cplx* r, cplx* z, hipStream_t st);
// w = J (hz - diag z), hz = H z  (device; w may be hz)
void mlp_neumann(const MlpStates& s, double theta, const cplx* hz, const cplx* z, cplx* w,
                 hipStream_t st);

// Work buffers of mlp_pair_loss (grown on demand); coef holds the pairs'
// scalars (theta, dot products, loss coefficients) on the device.
struct MlpLossWork {
    DevBuf t, q, g, part, coef;
};

// Training loss of a batch of P pairs of H's basis, rows of pair p at
// [p*n, (p+1)*n).  With K = P_u (H - theta) P_u (the operator of the
// correction equation, P_u = 1 - u u^H) and t = P_u (z + s1 w + s2 w2):
//     L_p = 1 - |<b, K t>|^2 / |K t|^2,   b = K t* / |K t*|,
// the cosine between the learned and the accurate correction t* measured in
// K's norm (K t* = -r up to the solver's tolerance).  z, w, w2, b, u (unit) are
// complex FP64, theta (P doubles) on the device too.  Returns the summed
// loss (want_loss; else 0 and the call does not wait for the device); with
// go non-null also go = d(mean loss)/d(o) (FP32, P*n*2): two batched applies
// of H per call.  Every per-pair scalar stays on the device.
double mlp_pair_loss(const Hamiltonian& H, const cplx* z, const cplx* w, const cplx* w2, const cplx* b,
                     const cplx* u, const double* theta, int P, const float* o, float* go,
                     MlpLossWork& wk, hipStream_t st, bool want_loss = true);

}  // namespace qk
