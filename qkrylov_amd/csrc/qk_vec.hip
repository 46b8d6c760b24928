// FP64 Krylov vector kernels (see qk_vec.hpp).  256-thread workgroups,
// wave64 butterfly reductions (__shfl_xor), then a fixed-order LDS reduction
// over the 4 waves.  Reference operations replaced: vector_ops.hpp:14-137
// (dot, norm, scal, axpy), the MGS loops (lanczos.cpp:149-151,
// davidson.cpp:213-215), Ritz/residual axpy chains (lanczos.cpp:194-196,
// davidson.cpp:161-165, 187, 202-205) and davidson_precond (:171-175).
#include "qk_vec.hpp"

#include <algorithm>

namespace qk {

namespace {

constexpr int kT = 256;                 // threads per workgroup
constexpr u32 kCH = Segments::kChunk;   // elements per reduction chunk
constexpr int kPer = kCH / kT;          // elements per thread per chunk

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over the workgroup; result valid in thread 0.  Fixed order.
__device__ __forceinline__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = ((sh[0] + sh[1]) + sh[2]) + sh[3];
    return r;
}

__device__ __forceinline__ cplx cmuld(cplx a, cplx b) {
    return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// ---------------- single-vector reductions ----------------
__global__ void __launch_bounds__(kT) k_dot_part(const cplx* __restrict__ a,
                                                 const cplx* __restrict__ b, u64 n,
                                                 cplx* __restrict__ part) {
    __shared__ double sh[4];
    const u64 start = u64(blockIdx.x) * kCH;
    double re = 0.0, im = 0.0;
    for (int k = 0; k < kPer; ++k) {
        const u64 i = start + u64(k) * kT + threadIdx.x;
        if (i < n) {
            const cplx x = a[i], y = b[i];
            re += x.re * y.re + x.im * y.im;
            im += x.re * y.im - x.im * y.re;
        }
    }
    re = block_sum(re, sh);
    im = block_sum(im, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = cplx{re, im};
}

__global__ void __launch_bounds__(kT) k_norm_part(const cplx* __restrict__ a, u64 n,
                                                  double* __restrict__ part) {
    __shared__ double sh[4];
    const u64 start = u64(blockIdx.x) * kCH;
    double s = 0.0;
    for (int k = 0; k < kPer; ++k) {
        const u64 i = start + u64(k) * kT + threadIdx.x;
        if (i < n) {
            const cplx x = a[i];
            s += x.re * x.re + x.im * x.im;
        }
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// k_norm_part for k vectors at stride ld (vector j = blockIdx.y): the same
// chunks and sums per vector, so each norm equals norm2's bit for bit
__global__ void __launch_bounds__(kT) k_norms_part(const cplx* __restrict__ a, u64 n, u64 ld,
                                                   double* __restrict__ part) {
    __shared__ double sh[4];
    const cplx* x = a + u64(blockIdx.y) * ld;
    const u64 start = u64(blockIdx.x) * kCH;
    double s = 0.0;
    for (int k = 0; k < kPer; ++k) {
        const u64 i = start + u64(k) * kT + threadIdx.x;
        if (i < n) {
            const cplx v = x[i];
            s += v.re * v.re + v.im * v.im;
        }
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[u64(blockIdx.y) * gridDim.x + blockIdx.x] = s;
}

// out[j] = sum_{b < nb} part[j*nb + b]   (one workgroup per j)
__global__ void __launch_bounds__(kT) k_finish_c(const cplx* __restrict__ part, u32 nb,
                                                 cplx* __restrict__ out) {
    __shared__ double sh[4];
    const cplx* p = part + u64(blockIdx.x) * nb;
    double re = 0.0, im = 0.0;
    for (u32 b = threadIdx.x; b < nb; b += kT) {
        re += p[b].re;
        im += p[b].im;
    }
    re = block_sum(re, sh);
    im = block_sum(im, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = cplx{re, im};
}

__global__ void __launch_bounds__(kT) k_finish_r(const double* __restrict__ part, u32 nb,
                                                 double* __restrict__ out) {
    __shared__ double sh[4];
    const double* p = part + u64(blockIdx.x) * nb;
    double s = 0.0;
    for (u32 b = threadIdx.x; b < nb; b += kT) s += p[b];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// ---------------- elementwise ----------------
__global__ void k_scale(cplx* __restrict__ x, u64 n, cplx s) {
    const u64 i = u64(blockIdx.x) * kT + threadIdx.x;
    if (i < n) x[i] = cmuld(x[i], s);
}

__global__ void k_scale_copy(const cplx* __restrict__ x, cplx* __restrict__ y, u64 n, cplx s) {
    const u64 i = u64(blockIdx.x) * kT + threadIdx.x;
    if (i < n) y[i] = cmuld(x[i], s);
}

__global__ void k_scale_to(const cplx* __restrict__ x, cplx* __restrict__ y, u64 n,
                           const double* __restrict__ s) {
    const u64 i = u64(blockIdx.x) * kT + threadIdx.x;
    const double c = s[0];
    if (i < n) y[i] = cplx{x[i].re * c, x[i].im * c};
}

__global__ void k_inv_norm(const double* __restrict__ nrm2, double* __restrict__ s, int k,
                           double thresh) {
    const int j = int(blockIdx.x) * kT + int(threadIdx.x);
    if (j >= k) return;
    const double t = sqrt(nrm2[j]);
    s[j] = t > thresh ? 1.0 / t : 0.0;
}

__global__ void k_axpy(cplx a, const cplx* __restrict__ x, cplx* __restrict__ y, u64 n) {
    const u64 i = u64(blockIdx.x) * kT + threadIdx.x;
    if (i < n) {
        const cplx p = cmuld(a, x[i]);
        y[i] = cplx{y[i].re + p.re, y[i].im + p.im};
    }
}

__global__ void k_lanczos_sub(cplx* __restrict__ w, const cplx* __restrict__ v,
                              const cplx* __restrict__ u, u64 n, const double* __restrict__ alpha,
                              const double* __restrict__ beta) {
    const u64 i = u64(blockIdx.x) * kT + threadIdx.x;
    if (i >= n) return;
    const double a = alpha[0];
    cplx r = w[i];
    const cplx vv = v[i];
    r.re -= a * vv.re;
    r.im -= a * vv.im;
    if (u) {
        const double b = beta[0];
        const cplx uu = u[i];
        r.re -= b * uu.re;
        r.im -= b * uu.im;
    }
    w[i] = r;
}

// h partials: part[j*nb + blk] = sum_{rows of blk} conj(V_j) w
// PER elements per thread: chunks of PER * kT entries, fewer per block on
// short vectors so that the grid still covers the chip
template <int PER>
__global__ void __launch_bounds__(kT) k_mdot_part(const cplx* const* __restrict__ V, int k,
                                                  const cplx* __restrict__ w, u64 n,
                                                  cplx* __restrict__ part, u32 nb) {
    __shared__ double sh[4];
    const u64 start = u64(blockIdx.x) * (PER * kT);
    cplx wl[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const u64 i = start + u64(e) * kT + threadIdx.x;
        wl[e] = (i < n) ? w[i] : cplx{0.0, 0.0};
    }
    for (int j = 0; j < k; ++j) {
        const cplx* vj = V[j];
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const u64 i = start + u64(e) * kT + threadIdx.x;
            if (i < n) {
                const cplx x = vj[i];
                re += x.re * wl[e].re + x.im * wl[e].im;
                im += x.re * wl[e].im - x.im * wl[e].re;
            }
        }
        re = block_sum(re, sh);
        im = block_sum(im, sh);
        if (threadIdx.x == 0) part[u64(j) * nb + blockIdx.x] = cplx{re, im};
    }
}

// w -= sum_j V_j h[j] over one chunk of kCH entries per block, with the
// block's partial of |w|^2 after the update when part is non-null (fixed
// order, deterministic).  Vectors outer, the thread's kPer entries inner: the
// kPer loads of one vector are independent and in flight together.
template <int PER>
__global__ void __launch_bounds__(kT) k_msub_chunk(const cplx* const* __restrict__ V, int k,
                                                   const cplx* __restrict__ h,
                                                   cplx* __restrict__ w, u64 n,
                                                   double* __restrict__ part) {
    __shared__ double sh[4];
    const u64 start = u64(blockIdx.x) * (PER * kT);
    const bool full = start + PER * kT <= n;
    cplx r[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const u64 i = start + u64(e) * kT + threadIdx.x;
        r[e] = (full || i < n) ? w[i] : cplx{0.0, 0.0};
    }
    for (int j = 0; j < k; ++j) {
        const cplx* vj = V[j];
        const cplx c = h[j];
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const u64 i = start + u64(e) * kT + threadIdx.x;
            if (full || i < n) {
                const cplx x = vj[i];
                r[e].re -= c.re * x.re - c.im * x.im;
                r[e].im -= c.re * x.im + c.im * x.re;
            }
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const u64 i = start + u64(e) * kT + threadIdx.x;
        if (full || i < n) {
            w[i] = r[e];
            acc += r[e].re * r[e].re + r[e].im * r[e].im;
        }
    }
    if (part) {
        acc = block_sum(acc, sh);
        if (threadIdx.x == 0) part[blockIdx.x] = acc;
    }
}

template <int NO>
__global__ void __launch_bounds__(kT) k_lincomb(const cplx* const* __restrict__ V, int m,
                                                const cplx* __restrict__ c, int nout,
                                                cplx* __restrict__ out, u64 ldo, u64 n) {
    const u64 i = u64(blockIdx.x) * kT + threadIdx.x;
    if (i >= n) return;
    cplx acc[NO];
#pragma unroll
    for (int o = 0; o < NO; ++o) acc[o] = cplx{0.0, 0.0};
    for (int j = 0; j < m; ++j) {
        const cplx x = V[j][i];
#pragma unroll
        for (int o = 0; o < NO; ++o) {
            if (o < nout) {
                const cplx cc = c[o * m + j];
                acc[o].re += cc.re * x.re - cc.im * x.im;
                acc[o].im += cc.re * x.im + cc.im * x.re;
            }
        }
    }
#pragma unroll
    for (int o = 0; o < NO; ++o)
        if (o < nout) out[u64(o) * ldo + i] = acc[o];
}

__global__ void k_jacobi(const cplx* __restrict__ r, const cplx* __restrict__ d, double lambda,
                         cplx* __restrict__ t, u64 n) {
    const u64 i = u64(blockIdx.x) * kT + threadIdx.x;
    if (i >= n) return;
    double dre = d[i].re - lambda, dim = d[i].im;
    if (sqrt(dre * dre + dim * dim) < 1e-10) {
        dre = (dre >= 0.0) ? 1e-10 : -1e-10;
        dim = 0.0;
    }
    const cplx x = r[i];
    if (dim == 0.0) {
        t[i] = cplx{x.re / dre, x.im / dre};
    } else {
        const double den = dre * dre + dim * dim;
        t[i] = cplx{(x.re * dre + x.im * dim) / den, (x.im * dre - x.re * dim) / den};
    }
}

// ---------------- segmented (batched) recurrences ----------------
__global__ void __launch_bounds__(kT) k_seg_alpha_part(
    const u64* __restrict__ coff, const u32* __restrict__ clen, const u32* __restrict__ csamp,
    const cplx* __restrict__ U, const cplx* __restrict__ Y, const int* __restrict__ alive,
    double* __restrict__ part) {
    __shared__ double sh[4];
    const u32 c = blockIdx.x;
    const u32 j = csamp[c];
    if (!alive[j]) {
        if (threadIdx.x == 0) part[c] = 0.0;
        return;
    }
    const u64 off = coff[c];
    const u32 len = clen[c];
    double s = 0.0;
    for (u32 e = threadIdx.x; e < len; e += kT) {
        const cplx u = U[off + e], y = Y[off + e];
        s += u.re * y.re + u.im * y.im;
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[c] = s;
}

__global__ void __launch_bounds__(kT) k_seg_alpha_fin(const u32* __restrict__ first,
                                                      const u32* __restrict__ count,
                                                      const double* __restrict__ part,
                                                      const double* __restrict__ cs,
                                                      double* __restrict__ alpha,
                                                      const int* __restrict__ alive) {
    __shared__ double sh[4];
    const u32 j = blockIdx.x;
    double s = 0.0;
    for (u32 b = threadIdx.x; b < count[j]; b += kT) s += part[first[j] + b];
    s = block_sum(s, sh);
    if (threadIdx.x == 0 && alive[j]) alpha[j] = cs[j] * cs[j] * s;
}

// alpha[j] = c_j^2 * sum_{b < count[j]} part[j*stride + b] for live samples
// (partials written by the apply kernel itself)
__global__ void __launch_bounds__(kT) k_alpha_from_parts(const double* __restrict__ part, u32 stride,
                                                         const u32* __restrict__ count,
                                                         const double* __restrict__ cs,
                                                         double* __restrict__ alpha,
                                                         const int* __restrict__ alive) {
    __shared__ double sh[4];
    const u32 j = blockIdx.x;
    double s = 0.0;
    for (u32 b = threadIdx.x; b < count[j]; b += kT) s += part[u64(j) * stride + b];
    s = block_sum(s, sh);
    if (threadIdx.x == 0 && alive[j]) alpha[j] = cs[j] * cs[j] * s;
}

__global__ void __launch_bounds__(kT) k_seg_update_part(
    const u64* __restrict__ coff, const u32* __restrict__ clen, const u32* __restrict__ csamp,
    cplx* __restrict__ Y, const cplx* __restrict__ U, const cplx* __restrict__ P,
    const double* __restrict__ cs, const double* __restrict__ cps,
    const double* __restrict__ bprev, const double* __restrict__ alpha,
    const int* __restrict__ alive, int first_step, double* __restrict__ part) {
    __shared__ double sh[4];
    const u32 c = blockIdx.x;
    const u32 j = csamp[c];
    if (!alive[j]) {
        if (threadIdx.x == 0) part[c] = 0.0;
        return;
    }
    const u64 off = coff[c];
    const u32 len = clen[c];
    const double cc = cs[j];
    const double ac = alpha[j] * cc;
    const double bp = first_step ? 0.0 : bprev[j] * cps[j];
    double s = 0.0;
    for (u32 e = threadIdx.x; e < len; e += kT) {
        const cplx y = Y[off + e], u = U[off + e];
        cplx w{cc * y.re - ac * u.re, cc * y.im - ac * u.im};
        if (!first_step) {
            const cplx p = P[off + e];
            w.re -= bp * p.re;
            w.im -= bp * p.im;
        }
        Y[off + e] = w;
        s += w.re * w.re + w.im * w.im;
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[c] = s;
}

__global__ void __launch_bounds__(kT) k_seg_update_fin(
    const u32* __restrict__ first, const u32* __restrict__ count, const double* __restrict__ part,
    double* __restrict__ cs, double* __restrict__ cps, double* __restrict__ bprev,
    double* __restrict__ beta, int* __restrict__ alive) {
    __shared__ double sh[4];
    const u32 j = blockIdx.x;
    double s = 0.0;
    for (u32 b = threadIdx.x; b < count[j]; b += kT) s += part[first[j] + b];
    s = block_sum(s, sh);
    if (threadIdx.x == 0 && alive[j]) {
        const double bt = sqrt(s);
        beta[j] = bt;
        cps[j] = cs[j];
        bprev[j] = bt;
        if (bt < 1e-15) {
            alive[j] = 0;
            cs[j] = 0.0;
        } else {
            cs[j] = 1.0 / bt;
        }
    }
}

__global__ void __launch_bounds__(kT) k_seg_norm_part(const u64* __restrict__ coff,
                                                      const u32* __restrict__ clen,
                                                      const cplx* __restrict__ X,
                                                      double* __restrict__ part) {
    __shared__ double sh[4];
    const u32 c = blockIdx.x;
    const u64 off = coff[c];
    const u32 len = clen[c];
    double s = 0.0;
    for (u32 e = threadIdx.x; e < len; e += kT) {
        const cplx x = X[off + e];
        s += x.re * x.re + x.im * x.im;
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[c] = s;
}

__global__ void __launch_bounds__(kT) k_seg_norm_fin(const u32* __restrict__ first,
                                                     const u32* __restrict__ count,
                                                     const double* __restrict__ part,
                                                     double* __restrict__ nrm) {
    __shared__ double sh[4];
    const u32 j = blockIdx.x;
    double s = 0.0;
    for (u32 b = threadIdx.x; b < count[j]; b += kT) s += part[first[j] + b];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) nrm[j] = sqrt(s);
}


// ---------------------------------------------------------------------------
// G = V^H W for m <= 64 vectors V and nb <= 16 vectors W (tall-skinny
// projection, Davidson's new rows of V^H H V) on FP64 MFMA.
// v_mfma_f64_16x16x4f64, lane l (r = l % 16, q = l / 16): A[r][q], B[q][r],
// D[q + 4i][r] for its four results i (measured on gfx950; the FP32/BF16
// 16x16 shapes use D[4q + i][r] instead).  The K index q covers two complex entries: q/2 selects the
// entry, q%2 the real or imaginary part, so
//   A = (Re v | Im v),  B = (Re w | Im w)    -> Re G = Re v Re w + Im v Im w
//   A = (Re v | Im v),  B' = (Im w | -Re w)  -> Im G = Re v Im w - Im v Re w.
// Every vector is read once per launch; per-workgroup partials are summed in
// a fixed order by k_block_dot_fin.
// ---------------------------------------------------------------------------
typedef double f64x4 __attribute__((ext_vector_type(4)));
constexpr int kBDMin = 128;   // least complex entries per wave slice

template <int MB>
__global__ void __launch_bounds__(kT)
k_block_dot_part(const cplx* const* __restrict__ V, const int m, const cplx* const* __restrict__ W,
                 const int nw, const u64 n, const u64 per_wave, cplx* __restrict__ part) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4, half = q >> 1, im = q & 1;
    const cplx* vp[MB];
#pragma unroll
    for (int b = 0; b < MB; ++b) vp[b] = (b * 16 + r < m) ? V[b * 16 + r] : nullptr;
    const cplx* wp = r < nw ? W[r] : nullptr;
    f64x4 acc_re[MB], acc_im[MB];
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        acc_re[b] = f64x4{0.0, 0.0, 0.0, 0.0};
        acc_im[b] = f64x4{0.0, 0.0, 0.0, 0.0};
    }
    const u64 w0 = (u64(blockIdx.x) * 4 + wave) * per_wave;
    const u64 w1 = min(n, w0 + per_wave);
    // 8 entries per round: lane half h takes entries c + 4h .. c + 4h + 3 in
    // the four MFMA steps, so each lane reads 64 contiguous bytes of each of
    // its vectors (L1 line reuse within the round instead of across rounds).
    // Loads past the slice or for absent vectors read a valid entry and are
    // replaced by zero, so no load sits under a branch.
    const bool wok = wp != nullptr;
    const cplx* wq = wok ? wp : W[0];
    bool vok[MB];
    const cplx* vq[MB];
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        vok[b] = vp[b] != nullptr;
        vq[b] = vok[b] ? vp[b] : V[0];
    }
    for (u64 c = w0; c < w1; c += 8) {
        const u64 e0 = c + 4 * u64(half);
        double wr[4], wi[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const cplx t = wq[min(e0 + s, w1 - 1)];
            const bool ok = wok && e0 + s < w1;
            wr[s] = ok ? t.re : 0.0;
            wi[s] = ok ? t.im : 0.0;
        }
#pragma unroll
        for (int b = 0; b < MB; ++b) {
            double vr[4], vi[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const cplx t = vq[b][min(e0 + s, w1 - 1)];
                const bool ok = vok[b] && e0 + s < w1;
                vr[s] = ok ? t.re : 0.0;
                vi[s] = ok ? t.im : 0.0;
            }
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const double a = im ? vi[s] : vr[s];
                const double bre = im ? wi[s] : wr[s];
                const double bsw = im ? -wr[s] : wi[s];
                acc_re[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bre, acc_re[b], 0, 0, 0);
                acc_im[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bsw, acc_im[b], 0, 0, 0);
            }
        }
    }
    // reduce the 4 waves in LDS (fixed order), write this workgroup's partial
    __shared__ double sh[4][MB * 16 * 16 * 2];
#pragma unroll
    for (int b = 0; b < MB; ++b)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = q + 4 * i;
            const int o = ((b * 16 + row) * 16 + r) * 2;
            sh[wave][o] = acc_re[b][i];
            sh[wave][o + 1] = acc_im[b][i];
        }
    __syncthreads();
    cplx* dst = part + size_t(blockIdx.x) * (MB * 16 * 16);
    for (int e = threadIdx.x; e < MB * 16 * 16; e += kT) {
        const double re = ((sh[0][2 * e] + sh[1][2 * e]) + sh[2][2 * e]) + sh[3][2 * e];
        const double iv = ((sh[0][2 * e + 1] + sh[1][2 * e + 1]) + sh[2][2 * e + 1]) + sh[3][2 * e + 1];
        dst[e] = cplx{re, iv};
    }
}

// G[i*nw + j] = sum of the workgroup partials, one workgroup per entry:
// thread t sums parts t, t + kT, ... in order, then a fixed-order tree
__global__ void __launch_bounds__(kT)
k_block_dot_fin(const cplx* __restrict__ part, const int nparts, const int stride, const int m,
                const int nw, cplx* __restrict__ G) {
    const int e = blockIdx.x;
    const int i = e / nw, j = e % nw;
    double re = 0.0, iv = 0.0;
    for (int b = threadIdx.x; b < nparts; b += kT) {
        const cplx v = part[size_t(b) * stride + i * 16 + j];
        re += v.re;
        iv += v.im;
    }
    __shared__ double sh[4];
    re = block_sum(re, sh);
    iv = block_sum(iv, sh);
    if (threadIdx.x == 0) G[e] = cplx{re, iv};
}

// W_j -= sum_i V_i G[i*nw + j] for j < nw (each V_i read once for all W_j)
template <int NO>
__global__ void __launch_bounds__(kT) k_block_sub(const cplx* const* __restrict__ V, int m,
                                                  cplx* const* __restrict__ W, int nw,
                                                  const cplx* __restrict__ G, u64 n) {
    const u64 i = u64(blockIdx.x) * kT + threadIdx.x;
    if (i >= n) return;
    cplx acc[NO];
#pragma unroll
    for (int o = 0; o < NO; ++o) acc[o] = o < nw ? W[o][i] : cplx{0.0, 0.0};
    for (int j = 0; j < m; ++j) {
        const cplx x = V[j][i];
#pragma unroll
        for (int o = 0; o < NO; ++o) {
            if (o < nw) {
                const cplx c = G[j * nw + o];
                acc[o].re -= c.re * x.re - c.im * x.im;
                acc[o].im -= c.re * x.im + c.im * x.re;
            }
        }
    }
#pragma unroll
    for (int o = 0; o < NO; ++o)
        if (o < nw) W[o][i] = acc[o];
}

// ---------------- Lanczos step with device-resident indices ----------------
// The step index k (*kd) and every scalar of step j (rec[kRec*j + f]) live
// on the device, so a step launches the same kernels with the same
// arguments whatever k is: steps can be queued ahead of the host's stopping
// test or replayed from one captured graph.  Record fields:
//   0 alpha = Re<v_k, H v_k>
//   1 |w|^2 after the three-term subtraction (before reorthogonalisation)
//   2 |w|^2 after the first projection pass
//   3 |w|^2 after the second pass (when it runs)
//   4 beta = |w| at the end of the step
//   5 1 when the second pass ran
// Pass 1 folds the three-term subtraction into the projection: with
// h_j = <V_j, w - alpha v_k - beta_{k-1} v_{k-1}>, the update is
// w -= sum_j c_j V_j with c_k = h_k + alpha, c_{k-1} = h_{k-1} + beta_{k-1}.
// Pass 2 (classical Gram-Schmidt once more) runs only when pass 1 removed
// more than half of |w|^2 (the Daniel-Gragg-Kaufman-Stewart test); every
// workgroup reads the same two recorded values, so the test is uniform.
constexpr int kRec = 8;

__global__ void k_fill_u32(u32* __restrict__ p, u64 n, u32 v) {
    const u64 i = u64(blockIdx.x) * kT + threadIdx.x;
    if (i < n) p[i] = v;
}

__device__ __forceinline__ bool lz_second(const double* rec, u32 k) {
    return !(rec[kRec * k + 2] > 0.5 * rec[kRec * k + 1]);
}

// alpha from the apply kernel's per-workgroup partial sums of Re<x, Hx>
__global__ void __launch_bounds__(kT) k_lz_alpha_parts(const double* __restrict__ part, u32 np,
                                                       const u32* __restrict__ kd,
                                                       double* __restrict__ rec) {
    __shared__ double sh[4];
    double s = 0.0;
    for (u32 b = threadIdx.x; b < np; b += kT) s += part[b];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) rec[kRec * kd[0]] = s;
}

// alpha from k_dot_part's complex partials (real part)
__global__ void __launch_bounds__(kT) k_lz_alpha_dot(const cplx* __restrict__ part, u32 nb,
                                                     const u32* __restrict__ kd,
                                                     double* __restrict__ rec) {
    __shared__ double sh[4];
    double s = 0.0;
    for (u32 b = threadIdx.x; b < nb; b += kT) s += part[b].re;
    s = block_sum(s, sh);
    if (threadIdx.x == 0) rec[kRec * kd[0]] = s;
}

// projection partials of pass `pass` over V_0..V_k: part[j*nb + blk]; pass 1
// also the block's partial of |w - alpha v_k - beta v_{k-1}|^2 in npart[blk]
template <int PER>
__global__ void __launch_bounds__(kT) k_lz_mdot(const cplx* const* __restrict__ V,
                                                const u32* __restrict__ kd,
                                                const cplx* __restrict__ w, u64 n,
                                                cplx* __restrict__ part, u32 nb,
                                                double* __restrict__ npart,
                                                const double* __restrict__ rec, int pass) {
    __shared__ double sh[4];
    const u32 k = kd[0];
    if (pass == 2 && !lz_second(rec, k)) return;
    const u64 start = u64(blockIdx.x) * (PER * kT);
    cplx wl[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const u64 i = start + u64(e) * kT + threadIdx.x;
        wl[e] = (i < n) ? w[i] : cplx{0.0, 0.0};
    }
    if (pass == 1) {
        const double a = rec[kRec * k];
        const cplx* vk = V[k];
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const u64 i = start + u64(e) * kT + threadIdx.x;
            if (i < n) {
                const cplx x = vk[i];
                wl[e].re -= a * x.re;
                wl[e].im -= a * x.im;
            }
        }
        if (k > 0) {
            const double b = rec[kRec * (k - 1) + 4];
            const cplx* vp = V[k - 1];
#pragma unroll
            for (int e = 0; e < PER; ++e) {
                const u64 i = start + u64(e) * kT + threadIdx.x;
                if (i < n) {
                    const cplx x = vp[i];
                    wl[e].re -= b * x.re;
                    wl[e].im -= b * x.im;
                }
            }
        }
        double s = 0.0;
#pragma unroll
        for (int e = 0; e < PER; ++e) s += wl[e].re * wl[e].re + wl[e].im * wl[e].im;
        s = block_sum(s, sh);
        if (threadIdx.x == 0) npart[blockIdx.x] = s;
    }
    for (u32 j = 0; j <= k; ++j) {
        const cplx* vj = V[j];
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const u64 i = start + u64(e) * kT + threadIdx.x;
            if (i < n) {
                const cplx x = vj[i];
                re += x.re * wl[e].re + x.im * wl[e].im;
                im += x.re * wl[e].im - x.im * wl[e].re;
            }
        }
        re = block_sum(re, sh);
        im = block_sum(im, sh);
        if (threadIdx.x == 0) part[u64(j) * nb + blockIdx.x] = cplx{re, im};
    }
}

// coefficients of pass `pass`: workgroup j <= k sums the partials of h_j
// (pass 1 adds alpha at j = k and beta_{k-1} at j = k - 1); workgroup k + 1
// of pass 1 records |w|^2 before the projection; the others have no work
__global__ void __launch_bounds__(kT) k_lz_coef(const cplx* __restrict__ part, u32 nb,
                                                const double* __restrict__ npart,
                                                const u32* __restrict__ kd,
                                                double* __restrict__ rec, cplx* __restrict__ c,
                                                int pass) {
    __shared__ double sh[4];
    const u32 k = kd[0];
    const u32 j = blockIdx.x;
    if (j > k + 1 || (pass == 2 && (j > k || !lz_second(rec, k)))) return;
    if (j == k + 1) {
        double s = 0.0;
        for (u32 b = threadIdx.x; b < nb; b += kT) s += npart[b];
        s = block_sum(s, sh);
        if (threadIdx.x == 0) rec[kRec * k + 1] = s;
        return;
    }
    const cplx* p = part + u64(j) * nb;
    double re = 0.0, im = 0.0;
    for (u32 b = threadIdx.x; b < nb; b += kT) {
        re += p[b].re;
        im += p[b].im;
    }
    re = block_sum(re, sh);
    im = block_sum(im, sh);
    if (threadIdx.x == 0) {
        if (pass == 1 && j == k) re += rec[kRec * k];
        if (pass == 1 && k > 0 && j == k - 1) re += rec[kRec * (k - 1) + 4];
        c[j] = cplx{re, im};
    }
}

// w -= sum_{j <= k} c_j V_j with the block's partial of |w|^2 afterwards
template <int PER>
__global__ void __launch_bounds__(kT) k_lz_msub(const cplx* const* __restrict__ V,
                                                const u32* __restrict__ kd,
                                                const cplx* __restrict__ c,
                                                cplx* __restrict__ w, u64 n,
                                                double* __restrict__ part,
                                                const double* __restrict__ rec, int pass) {
    __shared__ double sh[4];
    const u32 k = kd[0];
    if (pass == 2 && !lz_second(rec, k)) return;
    const u64 start = u64(blockIdx.x) * (PER * kT);
    const bool full = start + PER * kT <= n;
    cplx r[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const u64 i = start + u64(e) * kT + threadIdx.x;
        r[e] = (full || i < n) ? w[i] : cplx{0.0, 0.0};
    }
    for (u32 j = 0; j <= k; ++j) {
        const cplx* vj = V[j];
        const cplx cj = c[j];
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const u64 i = start + u64(e) * kT + threadIdx.x;
            if (full || i < n) {
                const cplx x = vj[i];
                r[e].re -= cj.re * x.re - cj.im * x.im;
                r[e].im -= cj.re * x.im + cj.im * x.re;
            }
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const u64 i = start + u64(e) * kT + threadIdx.x;
        if (full || i < n) {
            w[i] = r[e];
            acc += r[e].re * r[e].re + r[e].im * r[e].im;
        }
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// |w|^2 after pass `pass` into the record (field 2 or 3)
__global__ void __launch_bounds__(kT) k_lz_norm(const double* __restrict__ part, u32 nb,
                                                const u32* __restrict__ kd,
                                                double* __restrict__ rec, int pass) {
    __shared__ double sh[4];
    const u32 k = kd[0];
    if (pass == 2 && !lz_second(rec, k)) return;
    double s = 0.0;
    for (u32 b = threadIdx.x; b < nb; b += kT) s += part[b];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) rec[kRec * k + 1 + pass] = s;
}

// beta = |w| (the second pass's norm when it ran); V_{k+1} = w / beta (0
// when beta < 1e-15, the reference's breakdown threshold, so steps queued
// past a breakdown stay finite) and the same into vcur when given
__global__ void __launch_bounds__(kT) k_lz_next(cplx* const* __restrict__ V, const u32* __restrict__ kd,
                                                u32 nslots, const cplx* __restrict__ w, u64 n,
                                                double* __restrict__ rec, cplx* __restrict__ vcur) {
    const u32 k = kd[0];
    const bool second = lz_second(rec, k);
    const double beta = sqrt(second ? rec[kRec * k + 3] : rec[kRec * k + 2]);
    const double s = beta < 1e-15 ? 0.0 : 1.0 / beta;
    const u64 i = u64(blockIdx.x) * kT + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        rec[kRec * k + 4] = beta;
        rec[kRec * k + 5] = second ? 1.0 : 0.0;
    }
    if (i >= n) return;
    const cplx x = w[i];
    const cplx y = cplx{x.re * s, x.im * s};
    if (k + 1 < nslots) V[k + 1][i] = y;
    if (vcur) vcur[i] = y;
}

__global__ void k_lz_advance(u32* kd) { kd[0] += 1; }

inline u32 nblk(u64 n) { return static_cast<u32>((n + kCH - 1) / kCH); }
inline u32 ngrid(u64 n) { return static_cast<u32>((n + kT - 1) / kT); }

}  // namespace

// entries per thread: the largest power of two up to kPer that still gives
// about 1024 blocks (short vectors keep the chip covered)
static int per_thread(u64 n) {
    int per = kPer;
    while (per > 1 && (n + u64(per) * kT - 1) / (u64(per) * kT) < 1024) per >>= 1;
    return per;
}

static void launch_msub(const cplx* const* V, int k, const cplx* h, cplx* w, u64 n, double* part,
                        u32* nb_out, hipStream_t st) {
    const int per = per_thread(n);
    const u64 chunk = u64(per) * kT;
    const u32 nb = std::max<u32>(1, u32((n + chunk - 1) / chunk));
    if (nb_out) *nb_out = nb;
    switch (per) {
        case 16: hipLaunchKernelGGL(k_msub_chunk<16>, dim3(nb), dim3(kT), 0, st, V, k, h, w, n, part); break;
        case 8: hipLaunchKernelGGL(k_msub_chunk<8>, dim3(nb), dim3(kT), 0, st, V, k, h, w, n, part); break;
        case 4: hipLaunchKernelGGL(k_msub_chunk<4>, dim3(nb), dim3(kT), 0, st, V, k, h, w, n, part); break;
        case 2: hipLaunchKernelGGL(k_msub_chunk<2>, dim3(nb), dim3(kT), 0, st, V, k, h, w, n, part); break;
        default: hipLaunchKernelGGL(k_msub_chunk<1>, dim3(nb), dim3(kT), 0, st, V, k, h, w, n, part); break;
    }
}

void Segments::build(const std::vector<u64>& lens, hipStream_t st) {
    nsamp = static_cast<int>(lens.size());
    len = lens;
    base.assign(lens.size(), 0);
    total = 0;
    std::vector<u64> coff;
    std::vector<u32> clen, csamp, first(nsamp), count(nsamp);
    for (int j = 0; j < nsamp; ++j) {
        base[j] = total;
        first[j] = static_cast<u32>(coff.size());
        for (u64 o = 0; o < lens[j]; o += kChunk) {
            coff.push_back(total + o);
            clen.push_back(static_cast<u32>(std::min<u64>(kChunk, lens[j] - o)));
            csamp.push_back(static_cast<u32>(j));
        }
        count[j] = static_cast<u32>(coff.size()) - first[j];
        total += lens[j];
    }
    nchunks = static_cast<u32>(coff.size());
    upload(d_chunk_off, coff.data(), coff.size(), st);
    upload(d_chunk_len, clen.data(), clen.size(), st);
    upload(d_chunk_sample, csamp.data(), csamp.size(), st);
    upload(d_first, first.data(), first.size(), st);
    upload(d_count, count.data(), count.size(), st);
    partials.resize(sizeof(double) * (nchunks ? nchunks : 1));
    QK_HIP(hipStreamSynchronize(st));
}

namespace vec {

void dot(const cplx* a, const cplx* b, u64 n, cplx* out, DevBuf& scratch, hipStream_t st) {
    const u32 nb = std::max<u32>(1, nblk(n));
    scratch.resize(sizeof(cplx) * nb);
    hipLaunchKernelGGL(k_dot_part, dim3(nb), dim3(kT), 0, st, a, b, n, scratch.get<cplx>());
    hipLaunchKernelGGL(k_finish_c, dim3(1), dim3(kT), 0, st, scratch.get<cplx>(), nb, out);
    QK_HIP(hipGetLastError());
}

void norm2(const cplx* a, u64 n, double* out, DevBuf& scratch, hipStream_t st) {
    const u32 nb = std::max<u32>(1, nblk(n));
    scratch.resize(sizeof(double) * nb);
    hipLaunchKernelGGL(k_norm_part, dim3(nb), dim3(kT), 0, st, a, n, scratch.get<double>());
    hipLaunchKernelGGL(k_finish_r, dim3(1), dim3(kT), 0, st, scratch.get<double>(), nb, out);
    QK_HIP(hipGetLastError());
}

void norms(const cplx* a, int k, u64 n, u64 ld, double* out, DevBuf& scratch, hipStream_t st) {
    if (k <= 0) return;
    const u32 nb = std::max<u32>(1, nblk(n));
    scratch.resize(sizeof(double) * nb * size_t(k));
    hipLaunchKernelGGL(k_norms_part, dim3(nb, k), dim3(kT), 0, st, a, n, ld, scratch.get<double>());
    hipLaunchKernelGGL(k_finish_r, dim3(k), dim3(kT), 0, st, scratch.get<double>(), nb, out);
    QK_HIP(hipGetLastError());
}

void scale_copy(const cplx* x, cplx* y, u64 n, cplx s, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_scale_copy, dim3(ngrid(n)), dim3(kT), 0, st, x, y, n, s);
    QK_HIP(hipGetLastError());
}

void scale(cplx* x, u64 n, cplx s, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_scale, dim3(ngrid(n)), dim3(kT), 0, st, x, n, s);
    QK_HIP(hipGetLastError());
}

void scale_to(const cplx* x, cplx* y, u64 n, const double* s_dev, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_scale_to, dim3(ngrid(n)), dim3(kT), 0, st, x, y, n, s_dev);
    QK_HIP(hipGetLastError());
}

void inv_norm_or_zero(const double* nrm2, double* s, int k, double thresh, hipStream_t st) {
    if (k <= 0) return;
    hipLaunchKernelGGL(k_inv_norm, dim3((k + kT - 1) / kT), dim3(kT), 0, st, nrm2, s, k, thresh);
    QK_HIP(hipGetLastError());
}

void axpy(cplx a, const cplx* x, cplx* y, u64 n, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_axpy, dim3(ngrid(n)), dim3(kT), 0, st, a, x, y, n);
    QK_HIP(hipGetLastError());
}

void lanczos_sub(cplx* w, const cplx* v, const cplx* u, u64 n, const double* alpha,
                 const double* beta, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_lanczos_sub, dim3(ngrid(n)), dim3(kT), 0, st, w, v, u, n, alpha, beta);
    QK_HIP(hipGetLastError());
}

void mdot(const cplx* const* V, int k, const cplx* w, u64 n, cplx* h, DevBuf& scratch,
          hipStream_t st) {
    if (k <= 0) return;
    const int per = per_thread(n);
    const u64 chunk = u64(per) * kT;
    const u32 nb = std::max<u32>(1, u32((n + chunk - 1) / chunk));
    scratch.resize(sizeof(cplx) * nb * size_t(k));
    cplx* part = scratch.get<cplx>();
    switch (per) {
        case 16: hipLaunchKernelGGL(k_mdot_part<16>, dim3(nb), dim3(kT), 0, st, V, k, w, n, part, nb); break;
        case 8: hipLaunchKernelGGL(k_mdot_part<8>, dim3(nb), dim3(kT), 0, st, V, k, w, n, part, nb); break;
        case 4: hipLaunchKernelGGL(k_mdot_part<4>, dim3(nb), dim3(kT), 0, st, V, k, w, n, part, nb); break;
        case 2: hipLaunchKernelGGL(k_mdot_part<2>, dim3(nb), dim3(kT), 0, st, V, k, w, n, part, nb); break;
        default: hipLaunchKernelGGL(k_mdot_part<1>, dim3(nb), dim3(kT), 0, st, V, k, w, n, part, nb); break;
    }
    hipLaunchKernelGGL(k_finish_c, dim3(k), dim3(kT), 0, st, scratch.get<cplx>(), nb, h);
    QK_HIP(hipGetLastError());
}

namespace {
// Decomposition of block_dot: at most 4096 waves, each over at least kBDMin
// entries (long vectors keep long slices, short ones still spread over the
// chip), whole rounds of 8 entries; 4 waves per workgroup partial.
struct BDShape {
    u64 per_wave;
    u32 parts;    // workgroups (partials)
    int mb;       // 16-row blocks of V
    int stride;   // cplx per partial
};
BDShape bd_shape(int m, u64 n) {
    BDShape s{};
    const u64 waves = std::max<u64>(1, std::min<u64>((n + kBDMin - 1) / kBDMin, 4096));
    s.per_wave = std::max<u64>(8, (((n + waves - 1) / waves) + 7) & ~u64(7));
    s.parts = std::max<u32>(1, u32((((n + s.per_wave - 1) / s.per_wave) + 3) / 4));
    s.mb = (m + 15) / 16;
    s.stride = s.mb * 16 * 16;
    return s;
}
}  // namespace

size_t block_dot_scratch_bytes(int m, u64 n) {
    const BDShape s = bd_shape(m, n);
    return sizeof(cplx) * size_t(s.parts) * size_t(s.stride);
}

void block_dot_into(const cplx* const* V, int m, const cplx* const* W, int nw, u64 n, cplx* G,
                    cplx* part, hipStream_t st) {
    if (m <= 0 || nw <= 0) return;
    if (m > 64 || nw > 16) fail("block_dot: at most 64 x 16 projections per call");
    const BDShape s = bd_shape(m, n);
    const u32 g = s.parts;
    switch (s.mb) {
        case 1: hipLaunchKernelGGL((k_block_dot_part<1>), dim3(g), dim3(kT), 0, st, V, m, W, nw, n, s.per_wave, part); break;
        case 2: hipLaunchKernelGGL((k_block_dot_part<2>), dim3(g), dim3(kT), 0, st, V, m, W, nw, n, s.per_wave, part); break;
        case 3: hipLaunchKernelGGL((k_block_dot_part<3>), dim3(g), dim3(kT), 0, st, V, m, W, nw, n, s.per_wave, part); break;
        default: hipLaunchKernelGGL((k_block_dot_part<4>), dim3(g), dim3(kT), 0, st, V, m, W, nw, n, s.per_wave, part); break;
    }
    hipLaunchKernelGGL(k_block_dot_fin, dim3(m * nw), dim3(kT), 0, st, part, int(g), s.stride, m, nw, G);
    QK_HIP(hipGetLastError());
}

void block_dot(const cplx* const* V, int m, const cplx* const* W, int nw, u64 n, cplx* G,
               DevBuf& scratch, hipStream_t st) {
    if (m <= 0 || nw <= 0) return;
    if (m > 64 || nw > 16) fail("block_dot: at most 64 x 16 projections per call");
    scratch.resize(block_dot_scratch_bytes(m, n));
    block_dot_into(V, m, W, nw, n, G, scratch.get<cplx>(), st);
}

void block_sub(const cplx* const* V, int m, cplx* const* W, int nw, const cplx* G, u64 n,
               hipStream_t st) {
    if (m <= 0 || nw <= 0 || !n) return;
    if (nw > 16) fail("block_sub: at most 16 vectors per call");
    if (nw <= 4)
        hipLaunchKernelGGL(k_block_sub<4>, dim3(ngrid(n)), dim3(kT), 0, st, V, m, W, nw, G, n);
    else if (nw <= 8)
        hipLaunchKernelGGL(k_block_sub<8>, dim3(ngrid(n)), dim3(kT), 0, st, V, m, W, nw, G, n);
    else
        hipLaunchKernelGGL(k_block_sub<16>, dim3(ngrid(n)), dim3(kT), 0, st, V, m, W, nw, G, n);
    QK_HIP(hipGetLastError());
}

void msub(const cplx* const* V, int k, const cplx* h, cplx* w, u64 n, hipStream_t st) {
    if (k <= 0 || !n) return;
    launch_msub(V, k, h, w, n, nullptr, nullptr, st);
    QK_HIP(hipGetLastError());
}

void msub_norm(const cplx* const* V, int k, const cplx* h, cplx* w, u64 n, double* out,
               DevBuf& scratch, hipStream_t st) {
    if (!n) return;
    const int per = per_thread(n);
    const u32 nbmax = std::max<u32>(1, u32((n + u64(per) * kT - 1) / (u64(per) * kT)));
    scratch.resize(sizeof(double) * nbmax);
    u32 nb = 0;
    launch_msub(V, k, h, w, n, scratch.get<double>(), &nb, st);
    hipLaunchKernelGGL(k_finish_r, dim3(1), dim3(kT), 0, st, scratch.get<double>(), nb, out);
    QK_HIP(hipGetLastError());
}

LzStep::LzStep(u64 n_, int nslots_) : n(n_), nslots(nslots_) {
    per = per_thread(n);
    const u64 chunk = u64(per) * kT;
    nb = std::max<u32>(1, u32((n + chunk - 1) / chunk));
    nbd = std::max<u32>(1, nblk(n));
    const size_t ns = size_t(nslots) + 1;
    state.resize(sizeof(u32) * 64 + sizeof(double) * kRec * ns + sizeof(cplx) * ns);
    hpart.resize(sizeof(cplx) * (size_t(nb) * ns + std::max(nb, nbd)));
    npart.resize(sizeof(double) * nb);
}

u32* LzStep::kdev() const { return state.get<u32>(); }
double* LzStep::rec() const { return reinterpret_cast<double*>(state.get<char>() + sizeof(u32) * 64); }
cplx* LzStep::coef() const {
    return reinterpret_cast<cplx*>(state.get<char>() + sizeof(u32) * 64 +
                                   sizeof(double) * kRec * (size_t(nslots) + 1));
}

void LzStep::reset(u32 k, hipStream_t st) {
    // a fill kernel, not hipMemsetAsync, so the whole step stays capturable
    const size_t words = state.bytes() / sizeof(u32);
    hipLaunchKernelGGL(k_fill_u32, dim3(u32((words + kT - 1) / kT)), dim3(kT), 0, st, state.get<u32>(),
                       words, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kT), 0, st, state.get<u32>(), u64(1), k);
    QK_HIP(hipGetLastError());
}

void LzStep::alpha_from_parts(const double* part, u32 np, hipStream_t st) const {
    hipLaunchKernelGGL(k_lz_alpha_parts, dim3(1), dim3(kT), 0, st, part, np, kdev(), rec());
    QK_HIP(hipGetLastError());
}

void LzStep::alpha_dot(const cplx* v, const cplx* w, hipStream_t st) const {
    cplx* part = reinterpret_cast<cplx*>(hpart.get<cplx>() + size_t(nb) * (size_t(nslots) + 1));
    hipLaunchKernelGGL(k_dot_part, dim3(nbd), dim3(kT), 0, st, v, w, n, part);
    hipLaunchKernelGGL(k_lz_alpha_dot, dim3(1), dim3(kT), 0, st, part, nbd, kdev(), rec());
    QK_HIP(hipGetLastError());
}

void LzStep::reorth(const cplx* const* V, cplx* w, int pass, int coef_blocks, hipStream_t st) const {
    cplx* part = hpart.get<cplx>();
    double* np = npart.get<double>();
    switch (per) {
#define QK_LZ_MDOT(P)                                                                              \
    case P:                                                                                        \
        hipLaunchKernelGGL(k_lz_mdot<P>, dim3(nb), dim3(kT), 0, st, V, kdev(), w, n, part, nb, np, \
                           rec(), pass);                                                          \
        break;
        QK_LZ_MDOT(16) QK_LZ_MDOT(8) QK_LZ_MDOT(4) QK_LZ_MDOT(2)
        default: hipLaunchKernelGGL(k_lz_mdot<1>, dim3(nb), dim3(kT), 0, st, V, kdev(), w, n, part, nb,
                                    np, rec(), pass);
#undef QK_LZ_MDOT
    }
    hipLaunchKernelGGL(k_lz_coef, dim3(u32(coef_blocks)), dim3(kT), 0, st, part, nb, np, kdev(), rec(),
                       coef(), pass);
    switch (per) {
#define QK_LZ_MSUB(P)                                                                                    \
    case P:                                                                                              \
        hipLaunchKernelGGL(k_lz_msub<P>, dim3(nb), dim3(kT), 0, st, V, kdev(), coef(), w, n, np, rec(), \
                           pass);                                                                        \
        break;
        QK_LZ_MSUB(16) QK_LZ_MSUB(8) QK_LZ_MSUB(4) QK_LZ_MSUB(2)
        default: hipLaunchKernelGGL(k_lz_msub<1>, dim3(nb), dim3(kT), 0, st, V, kdev(), coef(), w, n, np,
                                    rec(), pass);
#undef QK_LZ_MSUB
    }
    hipLaunchKernelGGL(k_lz_norm, dim3(1), dim3(kT), 0, st, np, nb, kdev(), rec(), pass);
    QK_HIP(hipGetLastError());
}

void LzStep::finish(cplx* const* V, const cplx* w, cplx* vcur, hipStream_t st) const {
    hipLaunchKernelGGL(k_lz_next, dim3(std::max<u32>(1, ngrid(n))), dim3(kT), 0, st, V, kdev(),
                       u32(nslots), w, n, rec(), vcur);
    hipLaunchKernelGGL(k_lz_advance, dim3(1), dim3(1), 0, st, kdev());
    QK_HIP(hipGetLastError());
}

void lincomb(const cplx* const* V, int k, const cplx* c, cplx* out, u64 n, hipStream_t st) {
    lincomb_many(V, k, c, 1, out, n, n, st);
}

void lincomb_many(const cplx* const* V, int m, const cplx* c, int nout, cplx* out, u64 ldo,
                  u64 n, hipStream_t st) {
    if (!n || nout <= 0) return;
    for (int o0 = 0; o0 < nout; o0 += 8) {
        const int no = std::min(8, nout - o0);
        const cplx* cc = c + size_t(o0) * m;
        cplx* oo = out + u64(o0) * ldo;
        if (no == 1)
            hipLaunchKernelGGL(k_lincomb<1>, dim3(ngrid(n)), dim3(kT), 0, st, V, m, cc, no, oo,
                               ldo, n);
        else if (no <= 4)
            hipLaunchKernelGGL(k_lincomb<4>, dim3(ngrid(n)), dim3(kT), 0, st, V, m, cc, no, oo,
                               ldo, n);
        else
            hipLaunchKernelGGL(k_lincomb<8>, dim3(ngrid(n)), dim3(kT), 0, st, V, m, cc, no, oo,
                               ldo, n);
    }
    QK_HIP(hipGetLastError());
}

void jacobi_precond(const cplx* r, const cplx* d, double lambda, cplx* t, u64 n, hipStream_t st) {
    if (!n) return;
    hipLaunchKernelGGL(k_jacobi, dim3(ngrid(n)), dim3(kT), 0, st, r, d, lambda, t, n);
    QK_HIP(hipGetLastError());
}

void seg_alpha(Segments& seg, const cplx* U, const cplx* Y, const double* c, double* alpha_out,
               const int* alive, hipStream_t st) {
    if (!seg.nchunks) return;
    hipLaunchKernelGGL(k_seg_alpha_part, dim3(seg.nchunks), dim3(kT), 0, st,
                       seg.d_chunk_off.get<u64>(), seg.d_chunk_len.get<u32>(),
                       seg.d_chunk_sample.get<u32>(), U, Y, alive, seg.partials.get<double>());
    hipLaunchKernelGGL(k_seg_alpha_fin, dim3(seg.nsamp), dim3(kT), 0, st, seg.d_first.get<u32>(),
                       seg.d_count.get<u32>(), seg.partials.get<double>(), c, alpha_out, alive);
    QK_HIP(hipGetLastError());
}

void alpha_from_parts(const double* part, u32 stride, const u32* count, int nsamp, const double* c,
                      double* alpha_out, const int* alive, hipStream_t st) {
    if (nsamp <= 0) return;
    hipLaunchKernelGGL(k_alpha_from_parts, dim3(nsamp), dim3(kT), 0, st, part, stride, count, c,
                       alpha_out, alive);
    QK_HIP(hipGetLastError());
}

void seg_update(Segments& seg, cplx* Y, const cplx* U, const cplx* P, double* c, double* cp,
                double* beta_prev, const double* alpha, double* beta_out, int* alive,
                int first_step, hipStream_t st) {
    if (!seg.nchunks) return;
    hipLaunchKernelGGL(k_seg_update_part, dim3(seg.nchunks), dim3(kT), 0, st,
                       seg.d_chunk_off.get<u64>(), seg.d_chunk_len.get<u32>(),
                       seg.d_chunk_sample.get<u32>(), Y, U, P, c, cp, beta_prev, alpha, alive,
                       first_step, seg.partials.get<double>());
    hipLaunchKernelGGL(k_seg_update_fin, dim3(seg.nsamp), dim3(kT), 0, st,
                       seg.d_first.get<u32>(), seg.d_count.get<u32>(),
                       seg.partials.get<double>(), c, cp, beta_prev, beta_out, alive);
    QK_HIP(hipGetLastError());
}

void seg_norm(Segments& seg, const cplx* X, double* nrm_out, hipStream_t st) {
    if (!seg.nchunks) return;
    hipLaunchKernelGGL(k_seg_norm_part, dim3(seg.nchunks), dim3(kT), 0, st,
                       seg.d_chunk_off.get<u64>(), seg.d_chunk_len.get<u32>(), X,
                       seg.partials.get<double>());
    hipLaunchKernelGGL(k_seg_norm_fin, dim3(seg.nsamp), dim3(kT), 0, st, seg.d_first.get<u32>(),
                       seg.d_count.get<u32>(), seg.partials.get<double>(), nrm_out);
    QK_HIP(hipGetLastError());
}

}  // namespace vec
}  // namespace qk
