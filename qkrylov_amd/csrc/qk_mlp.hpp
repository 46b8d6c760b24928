// Learned preconditioner for the Davidson correction equation (H - theta) t = -r.
//
// A three-layer MLP (32 -> 128 -> 128 -> 1, ReLU) reads per-state features
// (diagonal element, theta, connectivity, local occupations) and returns a
// per-state correction w_i to the Jacobi multiplier 1/(H_ii - theta):
//     t_i = (1/(H_ii - theta) + o_i) * r_i .
// With the last layer at zero the network is exactly the Jacobi method.
// Weights are kept in FP32 (master copy, Adam moments) and cast to BF16 for
// the kernels; every matrix product runs on gfx950 MFMA (BF16 inputs, FP32
// accumulation).  Training minimises, per (residual, accurate correction)
// pair, 1 - cos^2 between P_u t and the accurate correction t*, where P_u
// projects out the Ritz vector u (Davidson orthogonalises against it).
#pragma once

#include "qk_device.hpp"

#include <vector>

namespace qk {

struct Hamiltonian;

constexpr int kMlpIn = 32;     // input features (zero-padded)
constexpr int kMlpHid = 128;   // hidden width

// FP32 parameter block, offsets in floats.  W1[k][n] and W2[k][n] are stored
// input-major (row k = input feature / hidden unit of the previous layer).
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpIn * kMlpHid;
constexpr int kMlpW2 = kMlpB1 + kMlpHid;
constexpr int kMlpB2 = kMlpW2 + kMlpHid * kMlpHid;
constexpr int kMlpW3 = kMlpB2 + kMlpHid;
constexpr int kMlpB3 = kMlpW3 + kMlpHid;
constexpr int kMlpParams = kMlpB3 + 1;

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

    // o = net(X) for n rows of kMlpIn BF16 features (device).  With h1/h2
    // non-null the BF16 activations are kept for backward().
    void forward(const u16* X, u64 n, float* o, u16* h1, u16* h2, hipStream_t st) const;
    // grads (device, kMlpParams floats) = d(sum_m go[m] o[m]) / d(params),
    // from the activations of forward().
    void backward(const u16* X, const u16* h1, const u16* h2, const float* go, u64 n,
                  float* grads, hipStream_t st);
    void adam_step(const float* grads, float lr, hipStream_t st);

    // Host-buffer forward/backward on given features (gradient checks).
    void eval_host(const float* X, u64 n, const float* go, float* o, float* grads);

    // Per-state features for (H, theta) into X (device, n x kMlpIn BF16).
    // `stat` holds the per-state statistics of H (see MlpStates).
    hipStream_t stream() const { return st_; }

    DevBuf params;            // FP32 master
    DevBuf m1, m2;            // Adam moments
    DevBuf w1t, w2t, w2k, w3b;   // BF16: W1^T [n][k], W2^T [n][k], W2 [k][n], w3
    int adam_t = 0;

private:
    void cast();
    hipStream_t st_ = nullptr;
    DevBuf part_, wg_;       // backward partial sums
};

// Per-state data of one Hamiltonian that the features need, on the device.
struct MlpStates {
    DevBuf diag;   // double: Re H_ii
    DevBuf stat;   // float[4 + 1]: occupation fractions (4), connectivity
    u64 n = 0;
    int nsites = 1;
    void build(const Hamiltonian& H, hipStream_t st);
};

// X[i] = features of state i for theta (device).
void mlp_features(const MlpStates& s, double theta, u16* X, hipStream_t st);
// t = (jacobi(theta) + o) * r  (device; the Jacobi clamp is the reference's)
void mlp_combine(const MlpStates& s, double theta, const float* o, const cplx* r, cplx* t,
                 hipStream_t st);

// Loss and per-state gradient of a batch of P pairs that share one basis:
// rows of pair p are [p*n, (p+1)*n).  r, ts (accurate corrections), u (Ritz
// vectors) are complex FP64, theta per pair.  Returns the summed loss over
// the pairs and writes go = d(mean loss)/d(o) (FP32, P*n).
double mlp_pair_loss(const MlpStates& s, const cplx* r, const cplx* ts, const cplx* u,
                     const double* theta_dev, int P, const float* o, float* go, DevBuf& scratch,
                     hipStream_t st);

}  // namespace qk
