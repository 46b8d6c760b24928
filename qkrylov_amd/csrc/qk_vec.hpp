// FP64 Krylov vector kernels for gfx950.  Every reduction is two-stage with
// a fixed decomposition (per-block partials, then one block per output in a
// fixed tree), so results are deterministic run to run.  Scalars that feed
// later kernels stay on the device (pointer arguments), so a sequence of
// steps needs no host synchronisation.
#pragma once

#include "qk_device.hpp"

#include <vector>

namespace qk {

// A batch of independent vectors laid out back to back in one flat buffer:
// sample j occupies [base[j], base[j] + len[j]).  Reductions are split into
// chunks of at most kChunk elements; chunk c belongs to sample chunk_sample[c]
// and the chunks of one sample are contiguous.
struct Segments {
    static constexpr u32 kChunk = 4096;
    int nsamp = 0;
    u64 total = 0;
    std::vector<u64> base, len;
    // device
    DevBuf d_chunk_off, d_chunk_len, d_chunk_sample, d_first, d_count;
    u32 nchunks = 0;
    DevBuf partials;   // nchunks cplx

    void build(const std::vector<u64>& lens, hipStream_t st);
};

namespace vec {

// out[0] = sum conj(a) b over n elements (device scalar).
void dot(const cplx* a, const cplx* b, u64 n, cplx* out, DevBuf& scratch, hipStream_t st);
// out[0] = sum |a|^2.
void norm2(const cplx* a, u64 n, double* out, DevBuf& scratch, hipStream_t st);
// out[j] = sum |a_j|^2 for the k vectors a_j = a + j*ld (device; one launch
// pair, each value equal to norm2's)
void norms(const cplx* a, int k, u64 n, u64 ld, double* out, DevBuf& scratch, hipStream_t st);
// x *= s (host scalar)
void scale(cplx* x, u64 n, cplx s, hipStream_t st);
// y = s x  (host scalar)
void scale_copy(const cplx* x, cplx* y, u64 n, cplx s, hipStream_t st);
// y = s * x (s on device, real)
void scale_to(const cplx* x, cplx* y, u64 n, const double* s_dev, hipStream_t st);
// s[j] = 1/sqrt(nrm2[j]) where sqrt(nrm2[j]) > thresh, else 0 (j < k; device)
void inv_norm_or_zero(const double* nrm2, double* s, int k, double thresh, hipStream_t st);
// y += a x  (host scalar)
void axpy(cplx a, const cplx* x, cplx* y, u64 n, hipStream_t st);
// w -= alpha v + beta u   (alpha, beta real device scalars; u may be null)
void lanczos_sub(cplx* w, const cplx* v, const cplx* u, u64 n, const double* alpha,
                 const double* beta, hipStream_t st);
// h[j] = sum conj(V_j) w, j < k; V: device array of k vector pointers.
void mdot(const cplx* const* V, int k, const cplx* w, u64 n, cplx* h, DevBuf& scratch,
          hipStream_t st);
// G[i*nw + j] = <V_i, W_j> = sum conj(V_i) W_j for m <= 64, nw <= 16 vectors
// (FP64 MFMA; each vector read once; deterministic two-stage reduction).
void block_dot(const cplx* const* V, int m, const cplx* const* W, int nw, u64 n, cplx* G,
               DevBuf& scratch, hipStream_t st);
// The same with caller-provided partial storage of block_dot_scratch_bytes(m, n)
// bytes (every byte is written before it is read: no clearing needed).
size_t block_dot_scratch_bytes(int m, u64 n);
void block_dot_into(const cplx* const* V, int m, const cplx* const* W, int nw, u64 n, cplx* G,
                    cplx* part, hipStream_t st);
// W_j -= sum_i V_i G[i*nw + j] for nw <= 16 vectors (G on device); each V_i
// is read once for all W_j
void block_sub(const cplx* const* V, int m, cplx* const* W, int nw, const cplx* G, u64 n,
               hipStream_t st);
// w -= sum_j V_j h[j] and *out = |w|^2 after the update (device double), in
// one pass over the vectors
void msub_norm(const cplx* const* V, int k, const cplx* h, cplx* w, u64 n, double* out,
               DevBuf& scratch, hipStream_t st);
// w -= sum_j V_j h[j]
void msub(const cplx* const* V, int k, const cplx* h, cplx* w, u64 n, hipStream_t st);
// out = sum_j c[j] V_j  (c on device)
void lincomb(const cplx* const* V, int k, const cplx* c, cplx* out, u64 n, hipStream_t st);
// out_k = sum_j c[k*m + j] V_j for k < nout (c on device), out_k at out + k*ldo
void lincomb_many(const cplx* const* V, int m, const cplx* c, int nout, cplx* out, u64 ldo,
                  u64 n, hipStream_t st);
// t = r / (d - lambda) with |d - lambda| < 1e-10 clamped to +-1e-10
// (reference src/solvers/davidson.cpp:171-175); d: complex diagonal.
void jacobi_precond(const cplx* r, const cplx* d, double lambda, cplx* t, u64 n, hipStream_t st);

// ---- one Lanczos step with full reorthogonalisation, every index and
// scalar on the device (the step index and a record of kRecFields doubles
// per step: alpha, |w|^2 before / after each projection pass, beta, whether
// the second pass ran).  The kernels of a step take the same arguments for
// every step, so steps can be queued ahead of the host's stopping test or
// replayed from one captured graph.  V: device array of nslots vector
// pointers (V_0..V_k read at step k, V_{k+1} written). ----
struct LzStep {
    static constexpr int kRecFields = 8;
    LzStep(u64 n, int nslots);
    u32* kdev() const;       // the step index
    double* rec() const;     // rec[kRecFields * j + f]
    cplx* coef() const;      // projection coefficients of the current pass
    // zero every record and set the step index to k (fill kernels, capturable)
    void reset(u32 k, hipStream_t st);
    // alpha of the current step from the apply's np partial sums of Re<x, Hx>
    void alpha_from_parts(const double* part, u32 np, hipStream_t st) const;
    // alpha of the current step as Re<v, w>
    void alpha_dot(const cplx* v, const cplx* w, hipStream_t st) const;
    // projection pass 1 (three-term subtraction folded in) or 2 (runs only
    // when pass 1 removed more than half of |w|^2); coef_blocks >= k + 2
    void reorth(const cplx* const* V, cplx* w, int pass, int coef_blocks, hipStream_t st) const;
    // beta, V_{k+1} = w / beta (and vcur when non-null), then k += 1
    void finish(cplx* const* V, const cplx* w, cplx* vcur, hipStream_t st) const;

    u64 n;
    int nslots;
    int per = 1;
    u32 nb = 1, nbd = 1;
    DevBuf state, hpart, npart;
};

// ---- batched recurrences over Segments (FTLM / continued fractions) ----
// Buffers U, Y, P are flat (seg.total elements).  Per-sample device arrays:
// c (scale of the current vector), cp (scale of the previous vector),
// beta_prev, alpha_out, beta_out, alive (1 while the recurrence runs).
//
// alpha_j = c_j^2 Re<U_j, Y_j>
void seg_alpha(Segments& seg, const cplx* U, const cplx* Y, const double* c, double* alpha_out,
               const int* alive, hipStream_t st);
// alpha[j] = c_j^2 * (sum of count[j] partials at part + j*stride), live j only
void alpha_from_parts(const double* part, u32 stride, const u32* count, int nsamp, const double* c,
                      double* alpha_out, const int* alive, hipStream_t st);
// W = c Y - alpha c U - beta_prev cp P (written into Y), beta_j = ||W_j||;
// then cp <- c, c <- 1/beta (0 and alive = 0 when beta < 1e-15), beta_prev <- beta.
void seg_update(Segments& seg, cplx* Y, const cplx* U, const cplx* P, double* c, double* cp,
                double* beta_prev, const double* alpha, double* beta_out, int* alive,
                int first_step, hipStream_t st);
// nrm_j = ||X_j||  (X flat)
void seg_norm(Segments& seg, const cplx* X, double* nrm_out, hipStream_t st);

}  // namespace vec
}  // namespace qk
