// MLP preconditioner kernels on gfx950: BF16 MFMA (v_mfma_f32_16x16x32_bf16)
// with FP32 accumulation for the forward pass, the activation gradients and
// the weight gradients; FP32 Adam on a master copy; FP64 loss reductions.
//
// MFMA operand layout (16x16x32, wave64; lane l, g = l / 16, c = l % 16):
//   A (16 x 32):  lane holds A[c][8g .. 8g+7]
//   B (32 x 16):  lane holds B[8g .. 8g+7][c]
//   D (16 x 16):  lane holds D[4g .. 4g+3][c]
// so an operand is 8 consecutive BF16 along k, one 16-byte load, when the
// source is stored k-contiguous (X and activations row-major, weights
// transposed to [n][k]).
#include "qk_mlp.hpp"

#include "qk_hamiltonian.hpp"
#include "qk_threads.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <thread>
#include <unordered_map>

namespace qk {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = kMlpIn, kHid = kMlpHid;
constexpr int kIn1 = kIn + 8, kHid1 = kHid + 8;   // padded LDS row lengths
constexpr int kRows = 64;          // rows per workgroup tile (4 waves x 16)
constexpr int kThreads = 256;
// forward: 8 waves share one LDS copy of the weights (about 80 KB per
// workgroup, two workgroups per CU, four waves per SIMD)
constexpr int kFwdRows = 128;
constexpr int kFwdThreads = 512;

__device__ __forceinline__ u16 d_f2bf(float f) {
    u32 u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return u16(u >> 16);
}
__device__ __forceinline__ float d_bf2f(u16 b) { return __uint_as_float(u32(b) << 16); }

__device__ __forceinline__ f32x4 mfma(const u16x8 a, const u16x8 b, const f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}
__device__ __forceinline__ u16x8 ld8(const u16* p) { return *reinterpret_cast<const u16x8*>(p); }

// ---------------------------------------------------------------------------
// forward: o = W3^T relu(W2^T relu(W1^T x + b1) + b2) + b3, per row (2 outputs).
// Activations are rounded to BF16 (they feed the next MFMA); h2 is rounded
// too, so backward sees exactly what forward used.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kFwdThreads)
k_mlp_forward(const u16* __restrict__ X, const u64 n, const u16* __restrict__ w1t,
              const u16* __restrict__ w2t, const u16* __restrict__ w3b,
              const float* __restrict__ prm, float* __restrict__ out, u16* __restrict__ h1s,
              u16* __restrict__ h2s) {
    // LDS rows padded by 16 bytes: the 16 lanes of an operand read hit 16
    // distinct 4-bank groups instead of one bank (row strides of 64 and 256
    // bytes map every row to the same banks)
    __shared__ __align__(16) u16 s_w1t[kHid * kIn1];
    __shared__ __align__(16) u16 s_w2t[kHid * kHid1];
    __shared__ __align__(16) u16 s_h1[kFwdRows * kHid1];
    __shared__ float s_b1[kHid], s_b2[kHid], s_w3[kHid * kMlpOut];
    for (int i = threadIdx.x; i < kHid * kIn / 8; i += kFwdThreads) {
        const int r = i / (kIn / 8), c8 = i % (kIn / 8);
        *reinterpret_cast<u16x8*>(s_w1t + r * kIn1 + c8 * 8) = reinterpret_cast<const u16x8*>(w1t)[i];
    }
    for (int i = threadIdx.x; i < kHid * kHid / 8; i += kFwdThreads) {
        const int r = i / (kHid / 8), c8 = i % (kHid / 8);
        *reinterpret_cast<u16x8*>(s_w2t + r * kHid1 + c8 * 8) = reinterpret_cast<const u16x8*>(w2t)[i];
    }
    for (int i = threadIdx.x; i < kHid; i += kFwdThreads) {
        s_b1[i] = prm[kMlpB1 + i];
        s_b2[i] = prm[kMlpB2 + i];
    }
    for (int i = threadIdx.x; i < kHid * kMlpOut; i += kFwdThreads) s_w3[i] = d_bf2f(w3b[i]);
    const float b30 = prm[kMlpB3], b31 = prm[kMlpB3 + 1];
    __syncthreads();

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    u16* h1w = s_h1 + wave * 16 * kHid1;
    for (u64 tile = blockIdx.x; tile * kFwdRows < n; tile += gridDim.x) {
        const u64 m0 = tile * kFwdRows + u64(wave) * 16;
        // layer 1: [16 x 32] x [32 x 128]
        const u64 ra = m0 + c;
        u16x8 a = {0, 0, 0, 0, 0, 0, 0, 0};
        if (ra < n) a = ld8(X + ra * kIn + 8 * g);
        f32x4 acc[8];
#pragma unroll
        for (int nb = 0; nb < 8; ++nb) {
            acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
            acc[nb] = mfma(a, ld8(s_w1t + (nb * 16 + c) * kIn1 + 8 * g), acc[nb]);
        }
#pragma unroll
        for (int nb = 0; nb < 8; ++nb) {
            const int col = nb * 16 + c;
#pragma unroll
            for (int i = 0; i < 4; ++i)
                h1w[(4 * g + i) * kHid1 + col] = d_f2bf(fmaxf(acc[nb][i] + s_b1[col], 0.f));
        }
        __syncthreads();
        if (h1s) {
            for (int q = lane; q < 16 * kHid / 8; q += 64) {
                const int rr = q / (kHid / 8), cc = (q % (kHid / 8)) * 8;
                if (m0 + rr < n)
                    *reinterpret_cast<u16x8*>(h1s + (m0 + rr) * kHid + cc) =
                        *reinterpret_cast<const u16x8*>(h1w + rr * kHid1 + cc);
            }
        }
        // layer 2: [16 x 128] x [128 x 128]
#pragma unroll
        for (int nb = 0; nb < 8; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < kHid / 32; ++kb) {
            const u16x8 a2 = ld8(h1w + c * kHid1 + kb * 32 + 8 * g);
#pragma unroll
            for (int nb = 0; nb < 8; ++nb)
                acc[nb] = mfma(a2, ld8(s_w2t + (nb * 16 + c) * kHid1 + kb * 32 + 8 * g), acc[nb]);
        }
        // layer 3 (two outputs) and the saved h2
        float part0[4] = {0.f, 0.f, 0.f, 0.f}, part1[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int nb = 0; nb < 8; ++nb) {
            const int col = nb * 16 + c;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const u16 hb = d_f2bf(fmaxf(acc[nb][i] + s_b2[col], 0.f));
                const float h = d_bf2f(hb);
                part0[i] += h * s_w3[col * kMlpOut];
                part1[i] += h * s_w3[col * kMlpOut + 1];
                const u64 row = m0 + 4 * g + i;
                if (h2s && row < n) h2s[row * kHid + col] = hb;
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                part0[i] += __shfl_xor(part0[i], off);
                part1[i] += __shfl_xor(part1[i], off);
            }
        if (c == 0) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const u64 row = m0 + 4 * g + i;
                if (row < n) {
                    out[row * kMlpOut] = part0[i] + b30;
                    out[row * kMlpOut + 1] = part1[i] + b31;
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// backward, activation part.  gz2 = (W3 go) * [h2 > 0] and
// gz1 = (gz2 W2^T) * [h1 > 0], both kept as BF16 for the weight-gradient
// products; per-workgroup FP32 partials of gw3 = h2^T go and gb3 = sum go.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kThreads)
k_mlp_backward_act(const u16* __restrict__ h1s, const u16* __restrict__ h2s,
                   const float* __restrict__ go, const u64 n, const u16* __restrict__ w2k,
                   const u16* __restrict__ w3b, u16* __restrict__ gz2s, u16* __restrict__ gz1s,
                   float* __restrict__ part3) {
    __shared__ __align__(16) u16 s_w2k[kHid * kHid1];   // W2[k][n]: B operand of gz2 W2^T
    __shared__ float s_w3[kHid * kMlpOut];
    __shared__ float s_gw3[4][kHid * kMlpOut];
    __shared__ float s_gb3[4][kMlpOut];
    for (int i = threadIdx.x; i < kHid * kHid / 8; i += kThreads) {
        const int r = i / (kHid / 8), c8 = i % (kHid / 8);
        *reinterpret_cast<u16x8*>(s_w2k + r * kHid1 + c8 * 8) = reinterpret_cast<const u16x8*>(w2k)[i];
    }
    for (int i = threadIdx.x; i < kHid * kMlpOut; i += kThreads) s_w3[i] = d_bf2f(w3b[i]);
    __syncthreads();

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    // gw3 partial: lane accumulates columns 32kb + 8g + j of its row c
    float gw3[kHid / 32][8][kMlpOut];
#pragma unroll
    for (int kb = 0; kb < kHid / 32; ++kb)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            gw3[kb][j][0] = 0.f;
            gw3[kb][j][1] = 0.f;
        }
    float gb30 = 0.f, gb31 = 0.f;
    for (u64 tile = blockIdx.x; tile * kRows < n; tile += gridDim.x) {
        const u64 m0 = tile * kRows + u64(wave) * 16;
        const u64 ra = m0 + c;
        const bool rv = ra < n;
        const float go0 = rv ? go[ra * kMlpOut] : 0.f;
        const float go1 = rv ? go[ra * kMlpOut + 1] : 0.f;
        if (g == 0) {
            gb30 += go0;
            gb31 += go1;
        }
        f32x4 acc[8];
#pragma unroll
        for (int nb = 0; nb < 8; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < kHid / 32; ++kb) {
            const int k0 = kb * 32 + 8 * g;
            u16x8 h2v = {0, 0, 0, 0, 0, 0, 0, 0};
            if (rv) h2v = ld8(h2s + ra * kHid + k0);
            u16x8 a;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const float h = d_bf2f(h2v[j]);
                gw3[kb][j][0] += go0 * h;
                gw3[kb][j][1] += go1 * h;
                const float gz = go0 * s_w3[(k0 + j) * kMlpOut] + go1 * s_w3[(k0 + j) * kMlpOut + 1];
                a[j] = d_f2bf(h > 0.f ? gz : 0.f);
            }
            if (rv) *reinterpret_cast<u16x8*>(gz2s + ra * kHid + k0) = a;
            // gh1[m][k] = sum_n gz2[m][n] W2[k][n]
#pragma unroll
            for (int nb = 0; nb < 8; ++nb)
                acc[nb] = mfma(a, ld8(s_w2k + (nb * 16 + c) * kHid1 + k0), acc[nb]);
        }
#pragma unroll
        for (int nb = 0; nb < 8; ++nb) {
            const int col = nb * 16 + c;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const u64 row = m0 + 4 * g + i;
                if (row < n) {
                    const float h1 = d_bf2f(h1s[row * kHid + col]);
                    gz1s[row * kHid + col] = d_f2bf(h1 > 0.f ? acc[nb][i] : 0.f);
                }
            }
        }
    }
    // reduce gw3 over the 16 rows (lanes c) of each column group, then waves
#pragma unroll
    for (int kb = 0; kb < kHid / 32; ++kb)
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int o = 0; o < kMlpOut; ++o)
#pragma unroll
                for (int off = 1; off < 16; off <<= 1)
                    gw3[kb][j][o] += __shfl_xor(gw3[kb][j][o], off);
    for (int off = 1; off < 16; off <<= 1) {
        gb30 += __shfl_xor(gb30, off);
        gb31 += __shfl_xor(gb31, off);
    }
    if (c == 0) {
#pragma unroll
        for (int kb = 0; kb < kHid / 32; ++kb)
#pragma unroll
            for (int j = 0; j < 8; ++j)
#pragma unroll
                for (int o = 0; o < kMlpOut; ++o)
                    s_gw3[wave][(kb * 32 + 8 * g + j) * kMlpOut + o] = gw3[kb][j][o];
    }
    if (lane == 0) {
        s_gb3[wave][0] = gb30;
        s_gb3[wave][1] = gb31;
    }
    __syncthreads();
    // W3 then b3, as in the parameter block
    float* dst = part3 + size_t(blockIdx.x) * (kHid * kMlpOut + kMlpOut);
    for (int i = threadIdx.x; i < kHid * kMlpOut; i += kThreads)
        dst[i] = ((s_gw3[0][i] + s_gw3[1][i]) + s_gw3[2][i]) + s_gw3[3][i];
    if (threadIdx.x < kMlpOut)
        dst[kHid * kMlpOut + threadIdx.x] =
            ((s_gb3[0][threadIdx.x] + s_gb3[1][threadIdx.x]) + s_gb3[2][threadIdx.x]) + s_gb3[3][threadIdx.x];
}

// ---------------------------------------------------------------------------
// weight gradients: G[K x 128] = A^T D and colsum[128] = sum_m D[m][:], per
// workgroup over its rows, A (n x K) and D (n x 128) BF16 row-major.  Tiles
// of 32 rows are transposed into LDS so both operands are k-contiguous.
// ---------------------------------------------------------------------------
template <int K>
__global__ void __launch_bounds__(kThreads)
k_mlp_wgrad(const u16* __restrict__ A, const u16* __restrict__ D, const u64 n,
            float* __restrict__ part) {
    constexpr int TR = 32;                       // rows per tile = MFMA K
    constexpr int NBLK = (K / 16) * (kHid / 16); // 16x16 output blocks
    constexpr int PER = NBLK / 4;                // per wave
    __shared__ __align__(16) u16 s_at[K * TR];
    __shared__ __align__(16) u16 s_dt[kHid * TR];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    f32x4 acc[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
    float colsum = 0.f;   // thread t < 128 owns column t
    for (u64 tile = blockIdx.x; tile * TR < n; tile += gridDim.x) {
        const u64 m0 = tile * TR;
        __syncthreads();
        for (int q = threadIdx.x; q < TR * K / 8; q += kThreads) {
            const int rr = q / (K / 8), kk = (q % (K / 8)) * 8;
            u16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
            if (m0 + rr < n) v = ld8(A + (m0 + rr) * K + kk);
#pragma unroll
            for (int j = 0; j < 8; ++j) s_at[(kk + j) * TR + rr] = v[j];
        }
        for (int q = threadIdx.x; q < TR * kHid / 8; q += kThreads) {
            const int rr = q / (kHid / 8), kk = (q % (kHid / 8)) * 8;
            u16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
            if (m0 + rr < n) v = ld8(D + (m0 + rr) * kHid + kk);
#pragma unroll
            for (int j = 0; j < 8; ++j) s_dt[(kk + j) * TR + rr] = v[j];
        }
        __syncthreads();
        if (threadIdx.x < kHid) {
            float s = 0.f;
            for (int rr = 0; rr < TR; ++rr) s += d_bf2f(s_dt[threadIdx.x * TR + rr]);
            colsum += s;
        }
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int blk = wave * PER + q;
            const int kb = blk / (kHid / 16), nb = blk % (kHid / 16);
            acc[q] = mfma(ld8(s_at + (kb * 16 + c) * TR + 8 * g), ld8(s_dt + (nb * 16 + c) * TR + 8 * g),
                          acc[q]);
        }
    }
    float* dst = part + size_t(blockIdx.x) * (K * kHid + kHid);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int blk = wave * PER + q;
        const int kb = blk / (kHid / 16), nb = blk % (kHid / 16);
#pragma unroll
        for (int i = 0; i < 4; ++i) dst[(kb * 16 + 4 * g + i) * kHid + nb * 16 + c] = acc[q][i];
    }
    if (threadIdx.x < kHid) dst[K * kHid + threadIdx.x] = colsum;
}

// grads[off + j] = sum over workgroups b (in order) of part[b][j], j < len
__global__ void k_mlp_sum_parts(const float* __restrict__ part, int nparts, int stride, int len,
                                float* __restrict__ dst) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= len) return;
    float s = 0.f;
    for (int b = 0; b < nparts; ++b) s += part[size_t(b) * stride + j];
    dst[j] = s;
}

// first stage of a long sum of partials: chunk c = blockIdx.y adds partials
// [c*per, (c+1)*per) in order into tmp[c][j]; k_mlp_sum_parts then adds the
// chunks in order (fixed order, deterministic, many chains in flight)
__global__ void k_mlp_sum_chunks(const float* __restrict__ part, int nparts, int stride, int len,
                                 int per, float* __restrict__ tmp) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= len) return;
    const int b0 = int(blockIdx.y) * per, b1 = min(nparts, b0 + per);
    float s = 0.f;
    for (int b = b0; b < b1; ++b) s += part[size_t(b) * stride + j];
    tmp[size_t(blockIdx.y) * len + j] = s;
}

__global__ void k_mlp_adam(float* __restrict__ p, float* __restrict__ m1, float* __restrict__ m2,
                           const float* __restrict__ g, int n, float lr, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = 0.9f * m1[i] + 0.1f * gi;
    const float b = 0.999f * m2[i] + 0.001f * gi * gi;
    m1[i] = a;
    m2[i] = b;
    p[i] -= lr * (a / c1) / (sqrtf(b / c2) + 1e-8f);
}

// BF16 copies: W1^T[n][k], W2^T[n][k], W2[k][n], w3
__global__ void k_mlp_cast(const float* __restrict__ p, u16* __restrict__ w1t, u16* __restrict__ w2t,
                           u16* __restrict__ w2k, u16* __restrict__ w3b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kIn * kHid) {
        const int k = i / kHid, nn = i % kHid;
        w1t[nn * kIn + k] = d_f2bf(p[kMlpW1 + i]);
    }
    if (i < kHid * kHid) {
        const int k = i / kHid, nn = i % kHid;
        const u16 b = d_f2bf(p[kMlpW2 + i]);
        w2t[nn * kHid + k] = b;
        w2k[i] = b;
    }
    if (i < kHid * kMlpOut) w3b[i] = d_f2bf(p[kMlpW3 + i]);
}

int mlp_grid(u64 n, u64 rows_per_tile) {
    const u64 tiles = (n + rows_per_tile - 1) / rows_per_tile;
    return int(std::max<u64>(1, std::min<u64>(tiles, u64(device_cu_count()) * 2)));
}

// ---------------------------------------------------------------------------
// features, combine, loss
// ---------------------------------------------------------------------------
__device__ __forceinline__ double jac_den(double d, double theta) {
    // the reference's clamp of the Jacobi denominator (davidson.cpp:166-177)
    double den = d - theta;
    if (fabs(den) < 1e-10) den = den >= 0.0 ? 1e-10 : -1e-10;
    return den;
}

// one network input row from H_ii (d), the state's five statistics and theta
__device__ __forceinline__ void mlp_feature_row(const double d, const float* __restrict__ st5,
                                                const double theta, const double inv_sites,
                                                u16* __restrict__ Xrow) {
    const double den = jac_den(d, theta);
    float f[kIn];
#pragma unroll
    for (int k = 0; k < kIn; ++k) f[k] = 0.f;
    f[0] = float(tanh((d - theta) * 0.25));
    f[1] = float(fmin(fmax(1.0 / den, -4.0), 4.0) * 0.25);
    f[2] = float(theta * inv_sites);
    f[3] = float(d * inv_sites);
    f[4] = st5[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) f[5 + k] = st5[k];
    f[9] = float(tanh(d - theta));
    u16x8 v[kIn / 8];
#pragma unroll
    for (int q = 0; q < kIn / 8; ++q)
#pragma unroll
        for (int j = 0; j < 8; ++j) v[q][j] = d_f2bf(f[q * 8 + j]);
#pragma unroll
    for (int q = 0; q < kIn / 8; ++q) reinterpret_cast<u16x8*>(Xrow)[q] = v[q];
}

__global__ void k_mlp_features(const double* __restrict__ diag, const float* __restrict__ stat,
                               const u64 n, const double theta, const double inv_sites,
                               u16* __restrict__ X) {
    const u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    mlp_feature_row(diag[i], stat + i * 5, theta, inv_sites, X + i * kIn);
}

__global__ void k_mlp_jacobi(const double* __restrict__ diag, const u64 n, const double theta,
                             const cplx* __restrict__ r, cplx* __restrict__ z) {
    const u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double j = 1.0 / jac_den(diag[i], theta);
    z[i] = cplx{j * r[i].re, j * r[i].im};
}

// w may alias hz (element i reads hz[i] before writing w[i])
__global__ void k_mlp_neumann(const double* __restrict__ diag, const u64 n, const double theta,
                              const cplx* hz, const cplx* __restrict__ z, cplx* w) {
    const u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = diag[i];
    const double j = 1.0 / jac_den(d, theta);
    w[i] = cplx{j * (hz[i].re - d * z[i].re), j * (hz[i].im - d * z[i].im)};
}

// the two coefficients of a state from the network output (a, b)
__device__ __forceinline__ void mlp_coef(const float* o, u64 k, double& s1, double& s2, double& ta,
                                         double& tb) {
    ta = tanh(double(o[k * kMlpOut]));
    tb = tanh(double(o[k * kMlpOut + 1]));
    s1 = -0.5 + 0.5 * ta;
    s2 = 0.5 * tb;
}

__device__ __forceinline__ cplx mlp_term(double s1, double s2, cplx z, cplx w, cplx w2) {
    return cplx{z.re + s1 * w.re + s2 * w2.re, z.im + s1 * w.im + s2 * w2.im};
}

// t = z + s1 w + s2 w2 with the coefficients of output row i
__global__ void k_mlp_combine(const u64 n, const float* __restrict__ o, const cplx* __restrict__ z,
                              const cplx* __restrict__ w, const cplx* __restrict__ w2,
                              cplx* __restrict__ t) {
    const u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s1, s2, ta, tb;
    mlp_coef(o, i, s1, s2, ta, tb);
    t[i] = mlp_term(s1, s2, z[i], w[i], w2[i]);
}

// ---- batched over the corrections of one iteration (grid.y = j) ----------
__global__ void k_mlp_jacobi_many(const double* __restrict__ diag, const u64 n, const MlpBatch b,
                                  cplx* __restrict__ z) {
    const u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= n) return;
    const double f = 1.0 / jac_den(diag[i], b.theta[j]);
    const cplx r = b.r[j][i];
    z[u64(j) * n + i] = cplx{f * r.re, f * r.im};
}

__global__ void k_mlp_neumann_many(const double* __restrict__ diag, const u64 n, const MlpBatch b,
                                   const cplx* hz, const cplx* __restrict__ z, cplx* w) {
    const u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= n) return;
    const double d = diag[i];
    const double f = 1.0 / jac_den(d, b.theta[j]);
    const u64 k = u64(j) * n + i;
    const cplx h = hz[k], zz = z[k];
    w[k] = cplx{f * (h.re - d * zz.re), f * (h.im - d * zz.im)};
}

// the features of k_mlp_features for row i of correction j
__global__ void k_mlp_features_many(const double* __restrict__ diag, const float* __restrict__ stat,
                                    const u64 rows, const MlpBatch b, const double inv_sites,
                                    u16* __restrict__ X) {
    const u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= rows) return;
    mlp_feature_row(diag[i], stat + i * 5, b.theta[j], inv_sites, X + (u64(j) * rows + i) * kIn);
}

// t_j = z_j + s1 w_j + s2 w2_j; partial sums of <u_j, t_j> per workgroup
constexpr int kCombT = 256;
__global__ void __launch_bounds__(kCombT)
k_mlp_combine_many(const u64 n, const float* __restrict__ o, const u64 rows, const u32* __restrict__ cls,
                   const cplx* __restrict__ z, const cplx* __restrict__ w, const cplx* __restrict__ w2,
                   const MlpBatch b, double* __restrict__ part) {
    const u64 i = u64(blockIdx.x) * kCombT + threadIdx.x;
    const int j = blockIdx.y;
    double dr = 0.0, di = 0.0;
    if (i < n) {
        double s1, s2, ta, tb;
        mlp_coef(o + u64(j) * rows * kMlpOut, cls ? cls[i] : i, s1, s2, ta, tb);
        const u64 k = u64(j) * n + i;
        const cplx t = mlp_term(s1, s2, z[k], w[k], w2[k]);
        b.t[j][i] = t;
        if (b.u[j]) {
            const cplx uu = b.u[j][i];
            dr = uu.re * t.re + uu.im * t.im;
            di = uu.re * t.im - uu.im * t.re;
        }
    }
    __shared__ double sh[2][kCombT / 64];
    for (int off = 32; off > 0; off >>= 1) {
        dr += __shfl_xor(dr, off);
        di += __shfl_xor(di, off);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = dr;
        sh[1][threadIdx.x >> 6] = di;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, c = 0.0;
        for (int q = 0; q < kCombT / 64; ++q) {
            a += sh[0][q];
            c += sh[1][q];
        }
        part[(u64(j) * gridDim.x + blockIdx.x) * 2] = a;
        part[(u64(j) * gridDim.x + blockIdx.x) * 2 + 1] = c;
    }
}

// c_j = sum of correction j's partials (fixed order), one workgroup per j
__global__ void __launch_bounds__(kCombT) k_mlp_dot_fin(const double* __restrict__ part, const int nblk,
                                                       cplx* __restrict__ c) {
    const int j = blockIdx.x;
    double a = 0.0, v = 0.0;
    for (int q = threadIdx.x; q < nblk; q += kCombT) {
        a += part[(size_t(j) * nblk + q) * 2];
        v += part[(size_t(j) * nblk + q) * 2 + 1];
    }
    __shared__ double sh[2][kCombT / 64];
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        v += __shfl_xor(v, off);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = a;
        sh[1][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0.0, y = 0.0;
        for (int q = 0; q < kCombT / 64; ++q) {
            x += sh[0][q];
            y += sh[1][q];
        }
        c[j] = cplx{x, y};
    }
}

__global__ void k_mlp_project_many(const u64 n, const MlpBatch b, const cplx* __restrict__ c) {
    const u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= n || !b.u[j]) return;
    const cplx cc = c[j], uu = b.u[j][i];
    cplx t = b.t[j][i];
    t.re -= cc.re * uu.re - cc.im * uu.im;
    t.im -= cc.re * uu.im + cc.im * uu.re;
    b.t[j][i] = t;
}

// ---- training loss ----------------------------------------------------------
// Per-pair reductions and updates over P pairs of n states stored pair-major
// (grid.y = pair).
constexpr int kLossT = 256;

template <int NS>
__device__ __forceinline__ void block_sums(double (&s)[NS], double* dst) {
    __shared__ double sh[NS][kLossT / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        double v = s[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) sh[q][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < NS) {
        double v = 0.0;
        for (int q = 0; q < kLossT / 64; ++q) v += sh[threadIdx.x][q];
        dst[threadIdx.x] = v;
    }
}

// per (pair, block): sum conj(x) y (re, im) and sum |y|^2
__global__ void __launch_bounds__(kLossT)
k_pair_dot(const cplx* __restrict__ x, const cplx* __restrict__ y, const u64 n, const int nblk,
           double* __restrict__ part) {
    const int p = blockIdx.y;
    double s[3] = {0, 0, 0};
    for (u64 i = u64(blockIdx.x) * kLossT + threadIdx.x; i < n; i += u64(nblk) * kLossT) {
        const u64 k = u64(p) * n + i;
        const cplx a = x[k], b = y[k];
        s[0] += a.re * b.re + a.im * b.im;
        s[1] += a.re * b.im - a.im * b.re;
        s[2] += b.re * b.re + b.im * b.im;
    }
    block_sums<3>(s, part + (size_t(p) * nblk + blockIdx.x) * 3);
}

// per pair: xy[p] = sum conj(x) y and (yy non-null) yy[p] = sum |y|^2 from
// k_pair_dot's partial sums, added in block order
__global__ void k_pair_dot_fin(const double* __restrict__ part, const int nblk, const int P,
                               cplx* __restrict__ xy, double* __restrict__ yy) {
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= P) return;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int q = 0; q < nblk; ++q) {
        const double* v = part + (size_t(p) * nblk + q) * 3;
        a += v[0];
        b += v[1];
        c += v[2];
    }
    xy[p] = cplx{a, b};
    if (yy) yy[p] = c;
}

// x_p -= c_p u_p
__global__ void k_pair_sub(const cplx* __restrict__ c, const cplx* __restrict__ u, cplx* __restrict__ x,
                           const u64 n, const int P) {
    const u64 k = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= u64(P) * n) return;
    const cplx cc = c[k / n], uu = u[k];
    x[k].re -= cc.re * uu.re - cc.im * uu.im;
    x[k].im -= cc.re * uu.im + cc.im * uu.re;
}

// y_p -= theta_p x_p
__global__ void k_pair_shift(const double* __restrict__ theta, const cplx* __restrict__ x,
                             cplx* __restrict__ y, const u64 n, const int P) {
    const u64 k = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= u64(P) * n) return;
    const double th = theta[k / n];
    y[k].re -= th * x[k].re;
    y[k].im -= th * x[k].im;
}

// per pair, from a = <b, q> and Q = |q|^2 (b unit): loss 1 - |a|^2 / Q and the
// coefficients of its gradient in q, g = ca b + cq q with ca = -2 a / Q,
// cq = 2 |a|^2 / Q^2
__global__ void k_pair_loss_coef(const cplx* __restrict__ a, const double* __restrict__ Q, const int P,
                                 double* __restrict__ loss, cplx* __restrict__ ca, double* __restrict__ cq) {
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= P) return;
    const double A2 = a[p].re * a[p].re + a[p].im * a[p].im;
    const double Qs = fmax(Q[p], 1e-300);
    loss[p] = 1.0 - A2 / Qs;
    ca[p] = cplx{-2.0 * a[p].re / Qs, -2.0 * a[p].im / Qs};
    cq[p] = 2.0 * A2 / (Qs * Qs);
}

// g_p = ca_p b_p + cq_p q_p
__global__ void k_pair_grad(const cplx* __restrict__ ca, const double* __restrict__ cq,
                            const cplx* __restrict__ b, const cplx* __restrict__ q, cplx* __restrict__ g,
                            const u64 n, const int P) {
    const u64 k = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= u64(P) * n) return;
    const int p = int(k / n);
    const cplx x = ca[p], bv = b[k], qv = q[k];
    const double y = cq[p];
    g[k] = cplx{x.re * bv.re - x.im * bv.im + y * qv.re, x.re * bv.im + x.im * bv.re + y * qv.im};
}

// go = d(loss / P)/d(a, b) from h = K g:  dL/ds1 = Re(conj(h) w), dL/ds2 = Re(conj(h) w2)
__global__ void k_mlp_grad_coef(const float* __restrict__ o, const cplx* __restrict__ w,
                                const cplx* __restrict__ w2, const cplx* __restrict__ h, const u64 cnt,
                                const double invP, float* __restrict__ go) {
    const u64 k = u64(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= cnt) return;
    double s1, s2, ta, tb;
    mlp_coef(o, k, s1, s2, ta, tb);
    const cplx hh = h[k], ww = w[k], vv = w2[k];
    const double g1 = hh.re * ww.re + hh.im * ww.im, g2 = hh.re * vv.re + hh.im * vv.im;
    go[k * kMlpOut] = float(g1 * 0.5 * (1.0 - ta * ta) * invP);
    go[k * kMlpOut + 1] = float(g2 * 0.5 * (1.0 - tb * tb) * invP);
}

}  // namespace

// ---------------------------------------------------------------------------
// Mlp
// ---------------------------------------------------------------------------
Mlp::Mlp(unsigned seed) {
    QK_HIP(hipStreamCreate(&st_));
    params.resize(sizeof(float) * kMlpParams);
    m1.resize(sizeof(float) * kMlpParams);
    m2.resize(sizeof(float) * kMlpParams);
    w1t.resize(sizeof(u16) * kIn * kHid);
    w2t.resize(sizeof(u16) * kHid * kHid);
    w2k.resize(sizeof(u16) * kHid * kHid);
    w3b.resize(sizeof(u16) * kHid * kMlpOut);
    // He-uniform hidden layers; the output layer starts at zero, so the
    // untrained network gives the damped Neumann correction z - w/2.
    std::vector<float> h(kMlpParams, 0.0f);
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    const float a1 = std::sqrt(6.0f / kIn), a2 = std::sqrt(6.0f / kHid);
    for (int i = 0; i < kIn * kHid; ++i) h[kMlpW1 + i] = a1 * u(rng);
    for (int i = 0; i < kHid * kHid; ++i) h[kMlpW2 + i] = a2 * u(rng);
    set_params(h.data());
}

Mlp::~Mlp() {
    if (st_) {
        (void)hipStreamSynchronize(st_);
        (void)hipStreamDestroy(st_);
    }
}

void Mlp::get_params(float* host) const {
    QK_HIP(hipMemcpyAsync(host, params.get(), sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, st_));
    QK_HIP(hipStreamSynchronize(st_));
}

void Mlp::set_params(const float* host) {
    QK_HIP(hipMemcpyAsync(params.get(), host, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, st_));
    QK_HIP(hipMemsetAsync(m1.get(), 0, sizeof(float) * kMlpParams, st_));
    QK_HIP(hipMemsetAsync(m2.get(), 0, sizeof(float) * kMlpParams, st_));
    adam_t = 0;
    cast();
    QK_HIP(hipStreamSynchronize(st_));
}

void Mlp::cast() {
    hipLaunchKernelGGL(k_mlp_cast, dim3((kHid * kHid + 255) / 256), dim3(256), 0, st_,
                       params.get<float>(), w1t.get<u16>(), w2t.get<u16>(), w2k.get<u16>(),
                       w3b.get<u16>());
    QK_HIP(hipGetLastError());
}

void Mlp::forward(const u16* X, u64 n, float* o, u16* h1, u16* h2, hipStream_t st) const {
    if (!n) return;
    hipLaunchKernelGGL(k_mlp_forward, dim3(mlp_grid(n, kFwdRows)), dim3(kFwdThreads), 0, st, X, n,
                       w1t.get<u16>(), w2t.get<u16>(), w3b.get<u16>(), params.get<float>(), o, h1, h2);
    QK_HIP(hipGetLastError());
}

void Mlp::backward(const u16* X, const u16* h1, const u16* h2, const float* go, u64 n,
                   float* grads, hipStream_t st) {
    QK_HIP(hipMemsetAsync(grads, 0, sizeof(float) * kMlpParams, st));
    if (!n) return;
    wg_.resize(sizeof(u16) * n * kHid * 2);
    u16* gz2 = wg_.get<u16>();
    u16* gz1 = gz2 + n * kHid;
    const int ga = mlp_grid(n, kRows);
    const int gw = mlp_grid(n, 32);
    const size_t s3 = size_t(kHid) * kMlpOut + kMlpOut, s1 = size_t(kIn) * kHid + kHid, s2 = size_t(kHid) * kHid + kHid;
    part_.resize(sizeof(float) * (size_t(ga) * s3 + size_t(gw) * (s1 + s2)));
    float* p3 = part_.get<float>();
    float* p1 = p3 + size_t(ga) * s3;
    float* p2 = p1 + size_t(gw) * s1;
    hipLaunchKernelGGL(k_mlp_backward_act, dim3(ga), dim3(kThreads), 0, st, h1, h2, go, n,
                       w2k.get<u16>(), w3b.get<u16>(), gz2, gz1, p3);
    hipLaunchKernelGGL((k_mlp_wgrad<kIn>), dim3(gw), dim3(kThreads), 0, st, X, gz1, n, p1);
    hipLaunchKernelGGL((k_mlp_wgrad<kHid>), dim3(gw), dim3(kThreads), 0, st, h1, gz2, n, p2);
    // W1 and b1 are contiguous in the parameter block, as are W2 and b2;
    // partials summed in chunks of kPer, then the chunks in order
    constexpr int kPer = 16;
    const int c1 = (gw + kPer - 1) / kPer, c3 = (ga + kPer - 1) / kPer;
    sum_.resize(sizeof(float) * (size_t(c1) * (s1 + s2) + size_t(c3) * s3));
    float* t1 = sum_.get<float>();
    float* t2 = t1 + size_t(c1) * s1;
    float* t3 = t2 + size_t(c1) * s2;
    auto two_stage = [&](const float* p, int np, size_t len, int nc, float* tmp, float* dst) {
        hipLaunchKernelGGL(k_mlp_sum_chunks, dim3((unsigned(len) + 255) / 256, unsigned(nc)), dim3(256), 0, st,
                           p, np, int(len), int(len), kPer, tmp);
        hipLaunchKernelGGL(k_mlp_sum_parts, dim3((int(len) + 255) / 256), dim3(256), 0, st, tmp, nc,
                           int(len), int(len), dst);
    };
    two_stage(p1, gw, s1, c1, t1, grads + kMlpW1);
    two_stage(p2, gw, s2, c1, t2, grads + kMlpW2);
    two_stage(p3, ga, s3, c3, t3, grads + kMlpW3);
    QK_HIP(hipGetLastError());
}

void Mlp::adam_step(const float* grads, float lr, hipStream_t st) {
    ++adam_t;
    const float c1 = 1.0f - std::pow(0.9f, float(adam_t));
    const float c2 = 1.0f - std::pow(0.999f, float(adam_t));
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, st,
                       params.get<float>(), m1.get<float>(), m2.get<float>(), grads, kMlpParams, lr,
                       c1, c2);
    hipLaunchKernelGGL(k_mlp_cast, dim3((kHid * kHid + 255) / 256), dim3(256), 0, st,
                       params.get<float>(), w1t.get<u16>(), w2t.get<u16>(), w2k.get<u16>(),
                       w3b.get<u16>());
    QK_HIP(hipGetLastError());
}

void Mlp::eval_host(const float* Xh, u64 n, const float* goh, float* oh, float* gradh) {
    std::vector<u16> xb(n * kIn);
    for (u64 i = 0; i < n * kIn; ++i) xb[i] = f32_to_bf16(Xh[i]);
    const u64 n1 = std::max<u64>(n, 1);
    DevBuf X(sizeof(u16) * n1 * kIn), o(sizeof(float) * n1 * kMlpOut), h1(sizeof(u16) * n1 * kHid),
        h2(sizeof(u16) * n1 * kHid), go(sizeof(float) * n1 * kMlpOut), g(sizeof(float) * kMlpParams);
    if (n) QK_HIP(hipMemcpyAsync(X.get(), xb.data(), sizeof(u16) * n * kIn, hipMemcpyHostToDevice, st_));
    forward(X.get<u16>(), n, o.get<float>(), h1.get<u16>(), h2.get<u16>(), st_);
    if (oh && n)
        QK_HIP(hipMemcpyAsync(oh, o.get(), sizeof(float) * n * kMlpOut, hipMemcpyDeviceToHost, st_));
    if (goh && gradh) {
        if (n)
            QK_HIP(hipMemcpyAsync(go.get(), goh, sizeof(float) * n * kMlpOut, hipMemcpyHostToDevice, st_));
        backward(X.get<u16>(), h1.get<u16>(), h2.get<u16>(), go.get<float>(), n, g.get<float>(), st_);
        QK_HIP(hipMemcpyAsync(gradh, g.get(), sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, st_));
    }
    QK_HIP(hipStreamSynchronize(st_));
}

// ---------------------------------------------------------------------------
// per-state statistics: occupation fractions and connectivity
// ---------------------------------------------------------------------------
std::shared_ptr<const MlpStates> mlp_states_of(const Hamiltonian& H) {
    if (!H.mlp_states) {
        auto s = std::make_shared<MlpStates>();
        s->build(H, H.stream);
        QK_HIP(hipStreamSynchronize(H.stream));
        H.mlp_states = std::move(s);
    }
    return H.mlp_states;
}

void MlpStates::build(const Hamiltonian& H, hipStream_t st) {
    const Basis& B = *H.basis;
    n = B.size();
    nsites = std::max(1, B.nsites());
    diag.resize(sizeof(double) * std::max<u64>(n, 1));
    stat.resize(sizeof(float) * 5 * std::max<u64>(n, 1));
    if (!n) return;
    DevBuf dc(sizeof(cplx) * n);
    H.diagonal(dc.get<cplx>(), st);
    std::vector<cplx> dh(n);
    QK_HIP(hipMemcpyAsync(dh.data(), dc.get(), sizeof(cplx) * n, hipMemcpyDeviceToHost, st));
    QK_HIP(hipStreamSynchronize(st));
    std::vector<double> dr(n);
    for (u64 i = 0; i < n; ++i) dr[i] = dh[i].re;

    // connectivity: fraction of off-diagonal groups that reach a basis state
    // with a non-zero element (the row's off-diagonal pattern)
    std::vector<const Group*> off;
    for (const Group& g : H.prog.groups)
        if (g.kind != kDiag) off.push_back(&g);
    const bool two = B.kind() == SiteKind::Hubbard || B.kind() == SiteKind::TJ;
    const int ns = B.nsites();
    std::vector<float> sh(5 * n);
    run_ranges(n, host_threads(n), [&](u64 lo, u64 hi) {
            for (u64 i = lo; i < hi; ++i) {
                const u64 s = B.state(i);
                float occ[4] = {0, 0, 0, 0};
                for (int q = 0; q < ns; ++q) {
                    const int v = two ? int((s >> (2 * q)) & 3u) : int((s >> q) & 1u);
                    occ[v] += 1.0f;
                }
                int cnt = 0;
                for (const Group* g : off) {
                    u32 cfg = 0;
                    for (size_t k = 0; k < g->pos.size(); ++k) cfg |= u32((s >> g->pos[k]) & 1u) << k;
                    const cplx a = g->table[cfg];
                    if (a.re == 0.0 && a.im == 0.0) continue;
                    if (B.index(s ^ g->flip) >= 0) ++cnt;
                }
                for (int k = 0; k < 4; ++k) sh[5 * i + k] = occ[k] / float(ns);
                sh[5 * i + 4] = off.empty() ? 0.0f : float(cnt) / float(off.size());
            }
        });
    QK_HIP(hipMemcpyAsync(diag.get(), dr.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
    QK_HIP(hipMemcpyAsync(stat.get(), sh.data(), sizeof(float) * 5 * n, hipMemcpyHostToDevice, st));

    // classes of states with bit-identical network inputs
    struct Key {
        u64 d;
        u32 f[5];
        bool operator==(const Key& o) const { return d == o.d && std::memcmp(f, o.f, sizeof f) == 0; }
    };
    struct KeyHash {
        size_t operator()(const Key& k) const {
            u64 h = k.d * 0x9E3779B97F4A7C15ull;
            for (u32 v : k.f) h = (h ^ v) * 0x100000001B3ull;
            return size_t(h ^ (h >> 29));
        }
    };
    std::unordered_map<Key, u32, KeyHash> ids;
    std::vector<u32> cl(n);
    std::vector<double> rd;
    std::vector<float> rs;
    ncls = 0;
    const u64 limit = n / 4;
    bool ok = true;
    for (u64 i = 0; i < n && ok; ++i) {
        Key k;
        std::memcpy(&k.d, &dr[i], 8);
        std::memcpy(k.f, &sh[5 * i], sizeof k.f);
        auto it = ids.find(k);
        if (it == ids.end()) {
            if (ids.size() >= std::max<u64>(limit, 1)) {
                ok = false;
                break;
            }
            it = ids.emplace(k, u32(ids.size())).first;
            rd.push_back(dr[i]);
            rs.insert(rs.end(), &sh[5 * i], &sh[5 * i] + 5);
        }
        cl[i] = it->second;
    }
    if (ok && !ids.empty()) {
        ncls = ids.size();
        upload(cls, cl.data(), cl.size(), st);
        upload(rdiag, rd.data(), rd.size(), st);
        upload(rstat, rs.data(), rs.size(), st);
    }
    QK_HIP(hipStreamSynchronize(st));
}

void mlp_features(const MlpStates& s, double theta, u16* X, hipStream_t st) {
    if (!s.n) return;
    hipLaunchKernelGGL(k_mlp_features, dim3(unsigned((s.n + 255) / 256)), dim3(256), 0, st,
                       s.diag.get<double>(), s.stat.get<float>(), s.n, theta, 1.0 / s.nsites, X);
    QK_HIP(hipGetLastError());
}

void mlp_jacobi(const MlpStates& s, double theta, const cplx* r, cplx* z, hipStream_t st) {
    if (!s.n) return;
    hipLaunchKernelGGL(k_mlp_jacobi, dim3(unsigned((s.n + 255) / 256)), dim3(256), 0, st,
                       s.diag.get<double>(), s.n, theta, r, z);
    QK_HIP(hipGetLastError());
}

void mlp_neumann(const MlpStates& s, double theta, const cplx* hz, const cplx* z, cplx* w,
                 hipStream_t st) {
    if (!s.n) return;
    hipLaunchKernelGGL(k_mlp_neumann, dim3(unsigned((s.n + 255) / 256)), dim3(256), 0, st,
                       s.diag.get<double>(), s.n, theta, hz, z, w);
    QK_HIP(hipGetLastError());
}

void mlp_jacobi_many(const MlpStates& s, const MlpBatch& b, cplx* z, hipStream_t st) {
    if (!s.n || b.nt <= 0) return;
    hipLaunchKernelGGL(k_mlp_jacobi_many, dim3(unsigned((s.n + 255) / 256), b.nt), dim3(256), 0, st,
                       s.diag.get<double>(), s.n, b, z);
    QK_HIP(hipGetLastError());
}

void mlp_neumann_many(const MlpStates& s, const MlpBatch& b, const cplx* hz, const cplx* z, cplx* w,
                      hipStream_t st) {
    if (!s.n || b.nt <= 0) return;
    hipLaunchKernelGGL(k_mlp_neumann_many, dim3(unsigned((s.n + 255) / 256), b.nt), dim3(256), 0, st,
                       s.diag.get<double>(), s.n, b, hz, z, w);
    QK_HIP(hipGetLastError());
}

void mlp_features_many(const MlpStates& s, const MlpBatch& b, u16* X, hipStream_t st) {
    if (!s.n || b.nt <= 0) return;
    const u64 rows = s.ncls ? s.ncls : s.n;
    const double* d = s.ncls ? s.rdiag.get<double>() : s.diag.get<double>();
    const float* f = s.ncls ? s.rstat.get<float>() : s.stat.get<float>();
    hipLaunchKernelGGL(k_mlp_features_many, dim3(unsigned((rows + 255) / 256), b.nt), dim3(256), 0, st,
                       d, f, rows, b, 1.0 / s.nsites, X);
    QK_HIP(hipGetLastError());
}

void mlp_combine_project_many(const MlpStates& s, const MlpBatch& b, const float* o, const cplx* z,
                              const cplx* w, const cplx* w2, DevBuf& scratch, hipStream_t st) {
    if (!s.n || b.nt <= 0) return;
    const u64 rows = s.ncls ? s.ncls : s.n;
    const int nblk = int((s.n + kCombT - 1) / kCombT);
    scratch.resize(sizeof(double) * 2 * size_t(nblk) * b.nt + sizeof(cplx) * kMlpBatch);
    double* part = scratch.get<double>();
    cplx* c = reinterpret_cast<cplx*>(part + 2 * size_t(nblk) * b.nt);
    hipLaunchKernelGGL(k_mlp_combine_many, dim3(nblk, b.nt), dim3(kCombT), 0, st, s.n, o, rows,
                       s.ncls ? s.cls.get<u32>() : nullptr, z, w, w2, b, part);
    bool any_u = false;
    for (int j = 0; j < b.nt; ++j) any_u = any_u || b.u[j];
    if (any_u) {
        hipLaunchKernelGGL(k_mlp_dot_fin, dim3(b.nt), dim3(kCombT), 0, st, part, nblk, c);
        hipLaunchKernelGGL(k_mlp_project_many, dim3(unsigned((s.n + 255) / 256), b.nt), dim3(256), 0, st,
                           s.n, b, c);
    }
    QK_HIP(hipGetLastError());
}

namespace {

// Per-pair scalars of mlp_pair_loss, all on the device: theta, the dot
// products and the loss coefficients of the P pairs.
struct PairScalars {
    const double* theta;
    cplx* xy;
    double* yy;
    cplx* ca;
    double* cq;
    double* loss;
};

PairScalars pair_scalars(MlpLossWork& wk, int P) {
    wk.coef.resize((sizeof(double) * 3 + sizeof(cplx) * 2) * size_t(P));
    unsigned char* b = static_cast<unsigned char*>(wk.coef.get());
    PairScalars s;
    s.xy = reinterpret_cast<cplx*>(b);
    s.ca = s.xy + P;
    s.yy = reinterpret_cast<double*>(s.ca + P);
    s.cq = s.yy + P;
    s.loss = s.cq + P;
    s.theta = nullptr;
    return s;
}

unsigned pair_grid(u64 n, int P) { return unsigned((u64(P) * n + 255) / 256); }

// xy_p = <x_p, y_p> and (yy non-null) yy_p = |y_p|^2, on the device
void pair_dot(const cplx* x, const cplx* y, u64 n, int P, cplx* xy, double* yy, MlpLossWork& wk,
              hipStream_t st) {
    const int nblk = int(std::min<u64>(64, (n + kLossT - 1) / kLossT));
    wk.part.resize(sizeof(double) * size_t(P) * nblk * 3);
    hipLaunchKernelGGL(k_pair_dot, dim3(nblk, P), dim3(kLossT), 0, st, x, y, n, nblk,
                       wk.part.get<double>());
    hipLaunchKernelGGL(k_pair_dot_fin, dim3(unsigned((P + 63) / 64)), dim3(64), 0, st,
                       wk.part.get<double>(), nblk, P, xy, yy);
    QK_HIP(hipGetLastError());
}

// x_p -= <u_p, x_p> u_p
void pair_project(const cplx* u, cplx* x, u64 n, int P, const PairScalars& sc, MlpLossWork& wk,
                  hipStream_t st) {
    pair_dot(u, x, n, P, sc.xy, nullptr, wk, st);
    hipLaunchKernelGGL(k_pair_sub, dim3(pair_grid(n, P)), dim3(256), 0, st, sc.xy, u, x, n, P);
    QK_HIP(hipGetLastError());
}

// y = P_u (H - theta) x for x orthogonal to u (P pairs)
void pair_k(const Hamiltonian& H, const cplx* u, const cplx* x, cplx* y, int P, const PairScalars& sc,
            MlpLossWork& wk, hipStream_t st) {
    const u64 n = H.dim();
    H.apply(x, y, P, n, st);
    hipLaunchKernelGGL(k_pair_shift, dim3(pair_grid(n, P)), dim3(256), 0, st, sc.theta, x, y, n, P);
    QK_HIP(hipGetLastError());
    pair_project(u, y, n, P, sc, wk, st);
}

}  // namespace

double mlp_pair_loss(const Hamiltonian& H, const cplx* z, const cplx* w, const cplx* w2, const cplx* b,
                     const cplx* u, const double* theta, int P, const float* o, float* go,
                     MlpLossWork& wk, hipStream_t st, bool want_loss) {
    const u64 n = H.dim();
    if (!n || P <= 0) return 0.0;
    const u64 tot = u64(P) * n;
    wk.t.resize(sizeof(cplx) * tot);
    wk.q.resize(sizeof(cplx) * tot);
    wk.g.resize(sizeof(cplx) * tot);
    cplx* t = wk.t.get<cplx>();
    cplx* q = wk.q.get<cplx>();
    cplx* g = wk.g.get<cplx>();
    PairScalars sc = pair_scalars(wk, P);
    sc.theta = theta;
    // t = P_u (z + s1 w + s2 w2), q = K t
    hipLaunchKernelGGL(k_mlp_combine, dim3(unsigned((tot + 255) / 256)), dim3(256), 0, st, tot, o, z, w,
                       w2, t);
    QK_HIP(hipGetLastError());
    pair_project(u, t, n, P, sc, wk, st);
    pair_k(H, u, t, q, P, sc, wk, st);
    // a = <b, q>, Q = |q|^2 (b unit); the loss and its coefficients per pair
    pair_dot(b, q, n, P, sc.xy, sc.yy, wk, st);
    hipLaunchKernelGGL(k_pair_loss_coef, dim3(unsigned((P + 63) / 64)), dim3(64), 0, st, sc.xy, sc.yy, P,
                       sc.loss, sc.ca, sc.cq);
    QK_HIP(hipGetLastError());
    if (go) {
        // g = d loss / d q, projected; t = h = K g; go from h
        hipLaunchKernelGGL(k_pair_grad, dim3(pair_grid(n, P)), dim3(256), 0, st, sc.ca, sc.cq, b, q, g, n,
                           P);
        QK_HIP(hipGetLastError());
        pair_project(u, g, n, P, sc, wk, st);
        pair_k(H, u, g, t, P, sc, wk, st);
        hipLaunchKernelGGL(k_mlp_grad_coef, dim3(unsigned((tot + 255) / 256)), dim3(256), 0, st, o, w, w2,
                           t, tot, 1.0 / double(P), go);
        QK_HIP(hipGetLastError());
    }
    if (!want_loss) return 0.0;
    std::vector<double> L(P);
    QK_HIP(hipMemcpyAsync(L.data(), sc.loss, sizeof(double) * size_t(P), hipMemcpyDeviceToHost, st));
    QK_HIP(hipStreamSynchronize(st));
    double loss = 0.0;
    for (int p = 0; p < P; ++p) loss += L[p];
    return loss;
}

}  // namespace qk
