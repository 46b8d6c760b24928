// Two-pass exchange apply, y = H x, for Hamiltonians whose off-diagonal
// groups are all real two-bit exchanges (the exchange path of qk_apply.hip).
//
// The basis is a ragged 2-D array: row = high half `hi` of the state, column
// = rank of the low half inside its class.  Within one class (the rows whose
// high halves have one popcount) the states form a rows x columns matrix X,
// and the bonds split three ways: high-half bonds mix rows at a fixed column
// (X -> A X), low-half bonds mix columns inside a row (X -> X B^T), and bonds
// straddling the split move a state to another row and another class.  One
// pass cannot keep both neighbourhoods on chip: a row is up to 55 KB and a
// high-half bond of the N = 28 chain reaches rows a class (188 MB) apart.
// So the apply runs in two passes:
//
//   pass C (k_xcol): high-half bonds.  A tile is up to 32 columns of two
//     rows of one class, the tiles of one band of 32 columns over every row
//     of the class are one chunk of work for one XCD, so the band of x
//     (3432 rows x 32 columns x 16 B = 1.8 MB at N = 28) stays in that XCD's
//     4 MB L2 while every row of the band gathers its partners from it.  It
//     writes the sum of the high-half terms to y.
//   pass R (k_xrow): everything else.  A workgroup stages the states of a row
//     group -- row hi and, when a straddling bond flips high-half bit P, row
//     hi ^ P -- in LDS, then evaluates the diagonal, every low-half bond and
//     the straddling bonds through P from LDS; the other straddling bonds
//     read x from global memory (consecutive ranks: coalesced, and the
//     neighbouring row groups are processed by the same XCD at the same
//     time).  It adds pass C's partial sum from y, writes y once more, and
//     the workgroup's partial sum of Re<x, y>.
//
// HBM traffic per state: pass C reads x (16 B) and writes y (16 B); pass R
// reads x (16 B) and y (16 B) and writes y (16 B): 80 B against 32 B
// algorithmic, in place of the single-pass kernel's re-reads of x by every
// high-half bond.  Each state's terms are added in a fixed order (diagonal,
// low-half bonds, the high-half sum, straddling bonds), the same for single
// and block applies.  Reference operation replaced: the CSR gather of
// src/hamiltonian/matrix_free_hamiltonian.cpp:156-167, y[a] = sum over the
// row's entries of conj(amp) x[col] (:75-84).
#include "qk_hamiltonian.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace qk {

namespace {

constexpr int kXpThreads = 1024;   // threads per workgroup (both passes)
constexpr int kXpWaves = kXpThreads / 64;
constexpr int kXpXcd = 8;          // XCDs on MI355X (workgroup b runs on XCD b % 8)
constexpr u32 kLdsLimit = 160 * 1024;

__device__ __forceinline__ u32 config_of_xp(u64 s, const DevGroup& g) {
    u32 c = 0;
    u64 pw = g.posw;
    for (u32 k = 0; k < g.nsup; ++k, pw >>= 6) c |= u32((s >> u32(pw & 63)) & 1) << k;
    return c;
}

__device__ __forceinline__ void nt_store(cplx* p, const cplx v) {
    __builtin_nontemporal_store(v.re, &p->re);
    __builtin_nontemporal_store(v.im, &p->im);
}

// Chunk schedule shared by both passes: chunk k (a band of pass C, a run of
// row groups of pass R) belongs to XCD k % 8; the `workers` of one XCD
// (waves for pass C, workgroups for pass R) sweep its chunks in order.
struct ChunkIter {
    const u32* off;   // chunk k: items [off[k], off[k+1])
    u32 nchunks, k, w, W, i, end;
    __device__ __forceinline__ bool next() {
        while (k < nchunks) {
            i = off[k] + w;
            end = off[k + 1];
            k += kXpXcd;
            if (i < end) return true;
        }
        return false;
    }
};

// ---------------------------------------------------------------- pass C
// One workgroup per band: stage x over columns [c0, c0 + w) of every row of
// the class in LDS (row rho at s_X[rho * w ..]), then thread rho sums its
// row's high-half partners from LDS for the w columns and writes them to y.
constexpr int kXpBandW = 4;    // widest band
constexpr int kXpStageC = 8;   // staged entries per thread per round

__global__ void __launch_bounds__(kXpThreads)
k_xcol(const XPassParams p, const cplx* __restrict__ x, cplx* __restrict__ y) {
    extern __shared__ __align__(16) unsigned char smem[];
    double* s_amp = reinterpret_cast<double*>(smem);
    cplx* s_X = reinterpret_cast<cplx*>(smem + 256 * sizeof(double));
    for (u32 i = threadIdx.x; i < p.ncpamp; i += kXpThreads) s_amp[i] = p.cpamp[i];
    ChunkIter it{p.cchunk_off, p.ncchunks, blockIdx.x % kXpXcd, blockIdx.x / kXpXcd,
                 gridDim.x / kXpXcd, 0, 0};
    while (it.next()) {
        for (u32 bi = it.i; bi < it.end; bi += it.W) {
            const XCBand Bd = p.cbands[bi];
            const u32 w = Bd.w, nr = Bd.nrows, ne = nr * w;
            const XCRow* rows = p.crow + Bd.row0;
            __syncthreads();   // the previous band is done with s_X (and s_amp is in)
            for (u32 e0 = 0; e0 < ne; e0 += kXpStageC * kXpThreads) {
                cplx v[kXpStageC];
#pragma unroll
                for (int k = 0; k < kXpStageC; ++k) {
                    const u32 e = e0 + threadIdx.x + u32(k) * kXpThreads;
                    v[k] = cplx{0.0, 0.0};
                    if (e < ne) {
                        const u32 rho = e / w, j = e - rho * w;
                        v[k] = x[rows[rho].base + Bd.c0 + j];
                    }
                }
#pragma unroll
                for (int k = 0; k < kXpStageC; ++k) {
                    const u32 e = e0 + threadIdx.x + u32(k) * kXpThreads;
                    if (e < ne) s_X[e] = v[k];
                }
            }
            __syncthreads();
            for (u32 rho = threadIdx.x; rho < nr; rho += kXpThreads) {
                const XCRow R = rows[rho];
                cplx acc[kXpBandW];
#pragma unroll
                for (int j = 0; j < kXpBandW; ++j) acc[j] = cplx{0.0, 0.0};
                const u32* pp = p.cpart + R.pptr;
                u32 e = 0;
                for (; e + 2 <= R.pcnt; e += 2) {
                    const u32 P0 = pp[e], P1 = pp[e + 1];
                    const double a0 = s_amp[P0 >> 24], a1 = s_amp[P1 >> 24];
                    const cplx* r0 = s_X + (P0 & 0xFFFFFFu) * w;
                    const cplx* r1 = s_X + (P1 & 0xFFFFFFu) * w;
                    cplx v0[kXpBandW], v1[kXpBandW];
#pragma unroll
                    for (int j = 0; j < kXpBandW; ++j) {
                        if (u32(j) < w) {
                            v0[j] = r0[j];
                            v1[j] = r1[j];
                        }
                    }
#pragma unroll
                    for (int j = 0; j < kXpBandW; ++j) {
                        if (u32(j) < w) {
                            acc[j].re += a0 * v0[j].re;
                            acc[j].im += a0 * v0[j].im;
                            acc[j].re += a1 * v1[j].re;
                            acc[j].im += a1 * v1[j].im;
                        }
                    }
                }
                if (e < R.pcnt) {
                    const u32 P0 = pp[e];
                    const double a0 = s_amp[P0 >> 24];
                    const cplx* r0 = s_X + (P0 & 0xFFFFFFu) * w;
#pragma unroll
                    for (int j = 0; j < kXpBandW; ++j) {
                        if (u32(j) < w) {
                            const cplx v0 = r0[j];
                            acc[j].re += a0 * v0.re;
                            acc[j].im += a0 * v0.im;
                        }
                    }
                }
#pragma unroll
                for (int j = 0; j < kXpBandW; ++j)
                    if (u32(j) < w) nt_store(&y[R.base + Bd.c0 + j], acc[j]);
            }
        }
    }
}

// ---------------------------------------------------------------- pass R
constexpr int kXpStage = 8;   // states per thread of one row group (groups <= 8192 states)
constexpr int kXpRound = 4;   // states per thread evaluated together

__global__ void __launch_bounds__(kXpThreads)
k_xrow(const XPassParams p, const cplx* __restrict__ x, cplx* __restrict__ y) {
    extern __shared__ __align__(16) unsigned char smem[];
    cplx* s_tab = reinterpret_cast<cplx*>(smem);
    u16* s_jlo = reinterpret_cast<u16*>(smem + size_t(p.ndmix_tab) * sizeof(cplx));
    cplx* s_x = reinterpret_cast<cplx*>(smem + p.lds_x_off);
    for (u32 i = threadIdx.x; i < p.ndmix_tab; i += kXpThreads) s_tab[i] = p.dmix_tab[i];
    for (u32 i = threadIdx.x; i < p.nlo; i += kXpThreads) s_jlo[i] = p.jlo[i];

    const u32 hi_end = p.n_lo + p.n_hi;
    const u32 mix_end = hi_end + p.n_mix;
    double dot = 0.0;
    ChunkIter it{p.gchunk_off, p.ngchunks, blockIdx.x % kXpXcd, blockIdx.x / kXpXcd,
                 gridDim.x / kXpXcd, 0, 0};
    while (it.next()) {
        for (u32 g = it.i; g < it.end; g += it.W) {
            const XGroup G = p.groups[g];
            const u32 n0 = G.cnt[0], n = n0 + G.cnt[1];
            // stage: every load of the group first, then the LDS stores
            cplx xr[kXpStage];
#pragma unroll
            for (int k = 0; k < kXpStage; ++k) {
                const u32 f = threadIdx.x + u32(k) * kXpThreads;
                xr[k] = cplx{0.0, 0.0};
                if (f < n) xr[k] = f < n0 ? x[G.base[0] + f] : x[G.base[1] + (f - n0)];
            }
            __syncthreads();   // the previous group is done with s_x (and the tables are in)
#pragma unroll
            for (int k = 0; k < kXpStage; ++k) {
                const u32 f = threadIdx.x + u32(k) * kXpThreads;
                if (f < n) s_x[f] = xr[k];
            }
            __syncthreads();
#pragma unroll
            for (int k0 = 0; k0 < kXpStage; k0 += kXpRound) {
                if (threadIdx.x + u32(k0) * kXpThreads >= n) break;
                u32 lo[kXpRound], rr[kXpRound], rh[kXpRound];
                bool on[kXpRound];
                double dl[kXpRound];
                cplx yp[kXpRound], acc[kXpRound], mx[kXpRound];
                // per-state tables and pass C's partial sums, loads issued together
#pragma unroll
                for (int k = 0; k < kXpRound; ++k) {
                    const u32 f = threadIdx.x + u32(k0 + k) * kXpThreads;
                    on[k] = f < n;
                    rh[k] = f < n0 ? 0u : 1u;
                    rr[k] = f - (rh[k] ? n0 : 0u);
                    const u32 pos = G.pos0[rh[k]] + rr[k];
                    lo[k] = on[k] ? p.lo_list[pos] : 0u;
                    dl[k] = on[k] ? p.dlo[pos] : 0.0;
                    yp[k] = cplx{0.0, 0.0};
                    if (on[k]) yp[k] = y[G.base[rh[k]] + rr[k]];
                }
                // diagonal
#pragma unroll
                for (int k = 0; k < kXpRound; ++k) {
                    acc[k] = cplx{0.0, 0.0};
                    if (!on[k]) continue;
                    const u32 f = threadIdx.x + u32(k0 + k) * kXpThreads;
                    const cplx xa = s_x[f];
                    double dg = dl[k] + G.dh[rh[k]];
                    if (p.n_dmix) {
                        const u64 st = (u64(G.hi[rh[k]]) << p.L) | lo[k];
                        for (u32 q = 0; q < p.n_dmix; ++q) {
                            const DevGroup D = p.dmix[q];
                            dg += s_tab[D.tab + config_of_xp(st, D)].re;
                        }
                    }
                    acc[k] = cplx{dg * xa.re, dg * xa.im};
                }
                // low-half bonds from LDS: one bond for all the round's
                // states at a time, so their LDS reads are in flight together
                for (u32 b = 0; b < p.n_lo; ++b) {
                    const XBond X = p.bonds[b];
                    u32 t[kXpRound];
                    double amp[kXpRound];
#pragma unroll
                    for (int k = 0; k < kXpRound; ++k) {
                        const u32 vi = (lo[k] >> X.bi) & 1, vj = (lo[k] >> X.bj) & 1;
                        const bool act = on[k] && vi != vj;
                        amp[k] = act ? (vi ? X.a10 : X.a01) : 0.0;
                        // an exchange keeps the class, so the partner lies in
                        // this row; a lane without one reads its own entry
                        const u32 f = threadIdx.x + u32(k0 + k) * kXpThreads;
                        const u32 tr = s_jlo[lo[k] ^ X.mlo];
                        t[k] = act ? (rh[k] ? n0 : 0u) + tr : (on[k] ? f : 0u);
                    }
                    cplx xv[kXpRound];
#pragma unroll
                    for (int k = 0; k < kXpRound; ++k) xv[k] = s_x[t[k]];
#pragma unroll
                    for (int k = 0; k < kXpRound; ++k) {
                        acc[k].re += amp[k] * xv[k].re;
                        acc[k].im += amp[k] * xv[k].im;
                    }
                }
#pragma unroll
                for (int k = 0; k < kXpRound; ++k) {
                    acc[k].re += yp[k].re;   // the high-half bonds (pass C)
                    acc[k].im += yp[k].im;
                }
                // straddling bonds: the paired row from LDS, others from x
                for (u32 b = hi_end; b < mix_end; ++b) {
                    const XBond X = p.bonds[b];
                    const bool paired = X.mhi == p.pair_mhi;
                    double amp[kXpRound];
#pragma unroll
                    for (int k = 0; k < kXpRound; ++k) {
                        const u32 hi = G.hi[rh[k]];
                        const u32 vi = (lo[k] >> X.bi) & 1, vj = (hi >> (X.bj - p.L)) & 1;
                        const u64 b2 = p.mnb[size_t(G.vr[rh[k]]) * p.n_mix + (b - hi_end)];
                        const bool act = on[k] && vi != vj && b2 != ~u64(0);
                        amp[k] = act ? (vi ? X.a10 : X.a01) : 0.0;
                        const u32 t = s_jlo[lo[k] ^ X.mlo];
                        mx[k] = cplx{0.0, 0.0};
                        if (act) {
                            if (paired && G.cnt[1 - rh[k]] != 0) mx[k] = s_x[(rh[k] ? 0u : n0) + t];
                            else mx[k] = x[b2 + t];
                        }
                    }
#pragma unroll
                    for (int k = 0; k < kXpRound; ++k) {
                        acc[k].re += amp[k] * mx[k].re;
                        acc[k].im += amp[k] * mx[k].im;
                    }
                }
#pragma unroll
                for (int k = 0; k < kXpRound; ++k) {
                    if (!on[k]) continue;
                    const u32 f = threadIdx.x + u32(k0 + k) * kXpThreads;
                    const cplx xa = s_x[f];
                    nt_store(&y[G.base[rh[k]] + rr[k]], acc[k]);
                    dot += xa.re * acc[k].re + xa.im * acc[k].im;
                }
            }
        }
    }
    if (p.dot_part) {
        // this workgroup's Re<x, y>, reduced in a fixed order
        __shared__ double s_dot[kXpWaves];
        const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (int off = 32; off > 0; off >>= 1) dot += __shfl_xor(dot, off);
        __syncthreads();
        if (lane == 0) s_dot[wave] = dot;
        __syncthreads();
        if (threadIdx.x == 0) {
            double tsum = 0.0;
            for (int w = 0; w < kXpWaves; ++w) tsum += s_dot[w];
            p.dot_part[blockIdx.x] = tsum;
        }
    }
}

// chunk offsets: chunk k covers items [off[k], off[k+1])
std::vector<u32> chunk_offsets(const std::vector<u32>& sizes) {
    std::vector<u32> off(sizes.size() + 1, 0);
    for (size_t k = 0; k < sizes.size(); ++k) off[k + 1] = off[k] + sizes[k];
    return off;
}

}  // namespace

// Set-up of the two-pass apply (after setup_exchange): row groups, column
// tiles, grids.  Leaves xpass off when a row group does not fit in LDS.
void Hamiltonian::setup_xpass() {
    xpass = false;
    if (!xchg || dim() == 0) return;
    const Basis& B = *basis;
    const int L = B.L();
    const std::vector<u64>& jh = B.jhi();
    const u32 nlo_b = u32(xparams.n_lo), nhi_b = u32(xparams.n_hi), nmix_b = u32(xparams.n_mix);
    const size_t nrow = jh.size();
    auto slot = [&](u32 h) -> u32 {
        if (!xparams.vrow) return h;
        // records indexed by the row's number among rows that hold states
        return u32(std::lower_bound(B.valid_hi().begin(), B.valid_hi().end(), h) - B.valid_hi().begin());
    };
    auto cnt_of = [&](u32 h) -> u32 {
        if (h >= nrow || jh[h] == ~u64(0)) return 0u;
        return B.cls_cnt()[B.hi_cls()[h]];
    };
    const u32 tables = u32(xparams.ndmix_tab * sizeof(cplx) + ((xparams.nlo * sizeof(u16) + 15) / 16) * 16);
    const u32 avail = kLdsLimit - std::min(kLdsLimit, tables + 256);
    u32 max_row = 0;
    for (u32 h : B.valid_hi()) max_row = std::max(max_row, cnt_of(h));
    const u32 max_states = u32(kXpStage) * kXpThreads;   // states one row group may hold
    if (max_row == 0 || u64(max_row) * sizeof(cplx) > avail || max_row > max_states) return;

    // pair mask: the straddling bond whose high-half bit pairs rows within
    // the LDS budget, preferring the bit farthest from the split (its
    // partner rows are farthest apart in memory)
    u32 pair = 0;
    for (u32 m = 0; m < nmix_b; ++m) {
        const u32 mh = bonds_host[nlo_b + nhi_b + m].mhi;
        if (popc(mh) != 1 || mh <= pair) continue;
        u32 worst = 0;
        for (u32 h : B.valid_hi()) worst = std::max(worst, cnt_of(h) + cnt_of(h ^ mh));
        if (u64(worst) * sizeof(cplx) <= avail && worst <= max_states) pair = mh;
    }
    // row groups: rows in ascending order of (hi without the pair bit), the
    // pair's rows together; consecutive groups differ in the lowest high-half
    // bits, whose straddling partners they read
    std::vector<XGroup> groups;
    std::vector<u32> gsz;
    u32 max_group = 0;
    for (u32 h : B.valid_hi()) {
        if (pair && (h & pair) && cnt_of(h ^ pair)) continue;   // taken with its partner
        XGroup G{};
        const u32 h2 = pair ? (h ^ pair) : h;
        const bool two = pair && cnt_of(h2) > 0;
        const u32 rows[2] = {h, two ? h2 : 0u};
        for (int k = 0; k < (two ? 2 : 1); ++k) {
            const u32 hh = rows[k];
            const int c = B.hi_cls()[hh];
            G.base[k] = jh[hh];
            G.hi[k] = hh;
            G.pos0[k] = B.cls_off()[c];
            G.cnt[k] = B.cls_cnt()[c];
            G.vr[k] = slot(hh);
            G.dh[k] = dhi_host[hh];
        }
        max_group = std::max(max_group, G.cnt[0] + G.cnt[1]);
        groups.push_back(G);
    }
    // chunks of consecutive groups, about 2 per workgroup of an XCD per chunk
    const int cus = device_cu_count();
    const u32 tab_x = (tables + 15) / 16 * 16;
    const u32 lds = tab_x + max_group * u32(sizeof(cplx));
    if (lds > kLdsLimit) return;
    const int per_cu = std::max(1, std::min<int>(2, int(kLdsLimit / lds)));
    xp_grid_r = std::max(kXpXcd, cus * per_cu / kXpXcd * kXpXcd);
    const u32 per_chunk = u32(std::max(1, 2 * xp_grid_r / kXpXcd));
    for (size_t g0 = 0; g0 < groups.size(); g0 += per_chunk)
        gsz.push_back(u32(std::min<size_t>(per_chunk, groups.size() - g0)));
    const std::vector<u32> goff = chunk_offsets(gsz);

    // pass C: per class, bands of w columns over every row of the class
    // (w = 4, 2 or 1: the widest whose band fits in LDS), the class's rows
    // with their high-half partners as (row index in the class, amplitude
    // id); chunks of as many consecutive bands as an XCD has workgroups, so
    // bands that share cache lines run on one XCD at the same time
    const int cus_c = device_cu_count();
    xp_grid_c = std::max(kXpXcd, cus_c / kXpXcd * kXpXcd);
    const u32 wg_per_xcd = u32(xp_grid_c / kXpXcd);
    const u32 lds_c_budget = kLdsLimit - 256 * u32(sizeof(double)) - 256;
    std::vector<std::vector<u32>> cls_rows(B.cls_cnt().size());
    for (u32 h : B.valid_hi()) cls_rows[B.hi_cls()[h]].push_back(h);
    std::vector<XCBand> cbands;
    std::vector<XCRow> crow;
    std::vector<u32> cpart, csz;
    std::vector<double> cpamp;
    auto amp_id = [&](double a) -> u32 {
        for (size_t i = 0; i < cpamp.size(); ++i)
            if (cpamp[i] == a) return u32(i);
        cpamp.push_back(a);
        return u32(cpamp.size() - 1);
    };
    u32 max_band = 0;
    for (size_t c = 0; c < cls_rows.size(); ++c) {
        const std::vector<u32>& rs = cls_rows[c];
        const u32 width = B.cls_cnt()[c];
        if (rs.empty() || width == 0) continue;
        if (rs.size() > 0xFFFFFFu || u64(rs.size()) * sizeof(cplx) > lds_c_budget) return;
        u32 w = 4;
        while (w > 1 && u64(rs.size()) * w * sizeof(cplx) > lds_c_budget) w >>= 1;
        w = std::min(w, width);
        std::vector<u32> local(rs.size());
        std::vector<std::pair<u32, u32>> idx;   // (hi, local index) for lookups
        idx.reserve(rs.size());
        for (u32 i = 0; i < rs.size(); ++i) idx.emplace_back(rs[i], i);
        std::sort(idx.begin(), idx.end());
        auto local_of = [&](u32 h) -> u32 {
            auto itx = std::lower_bound(idx.begin(), idx.end(), std::make_pair(h, 0u));
            if (itx == idx.end() || itx->first != h) fail("two-pass apply: high-half partner row missing");
            return itx->second;
        };
        const u32 row0 = u32(crow.size());
        for (u32 h : rs) {
            XCRow R{};
            R.base = jh[h];
            R.pptr = u32(cpart.size());
            for (u32 b = 0; b < nhi_b; ++b) {
                const XBond& X = bonds_host[nlo_b + b];
                const u32 vi = (h >> (X.bi - L)) & 1u, vj = (h >> (X.bj - L)) & 1u;
                if (vi == vj) continue;
                const u32 id = amp_id(vi ? X.a10 : X.a01);
                if (id > 255) return;
                cpart.push_back(local_of(h ^ X.mhi) | (id << 24));
            }
            R.pcnt = u32(cpart.size()) - R.pptr;
            crow.push_back(R);
        }
        std::vector<XCBand> cb;
        for (u32 c0 = 0; c0 < width; c0 += w)
            cb.push_back(XCBand{row0, u32(rs.size()), c0, std::min(w, width - c0)});
        max_band = std::max(max_band, u32(rs.size()) * w);
        for (size_t b0 = 0; b0 < cb.size(); b0 += wg_per_xcd) {
            const size_t cnt = std::min<size_t>(wg_per_xcd, cb.size() - b0);
            cbands.insert(cbands.end(), cb.begin() + b0, cb.begin() + b0 + cnt);
            csz.push_back(u32(cnt));
        }
    }
    xp_lds_c = int(256 * sizeof(double) + max_band * sizeof(cplx));
    if (xp_lds_c > int(kLdsLimit)) return;
    const std::vector<u32> coff = chunk_offsets(csz);

    upload(d_xgroups, groups.data(), groups.size(), stream);
    upload(d_xgchunk, goff.data(), goff.size(), stream);
    upload(d_xcbands, cbands.data(), cbands.size(), stream);
    upload(d_xcrow, crow.data(), crow.size(), stream);
    upload(d_xcpart, cpart.data(), cpart.size(), stream);
    upload(d_xcpamp, cpamp.data(), cpamp.size(), stream);
    upload(d_xcchunk, coff.data(), coff.size(), stream);
    QK_HIP(hipStreamSynchronize(stream));

    XPassParams& q = xpp;
    q = XPassParams{};
    q.groups = d_xgroups.get<XGroup>();
    q.gchunk_off = d_xgchunk.get<u32>();
    q.ngchunks = u32(gsz.size());
    q.cbands = d_xcbands.get<XCBand>();
    q.crow = d_xcrow.get<XCRow>();
    q.cpart = d_xcpart.get<u32>();
    q.cpamp = d_xcpamp.get<double>();
    q.ncpamp = u32(cpamp.size());
    q.cchunk_off = d_xcchunk.get<u32>();
    q.ncchunks = u32(csz.size());
    q.jlo = xparams.jlo;
    q.lo_list = xparams.lo_list;
    q.dlo = xparams.dlo;
    q.bonds = xparams.bonds;
    q.dmix = xparams.dmix;
    q.dmix_tab = xparams.dmix_tab;
    q.hnb = xparams.hnb;
    q.hcnt = xparams.hcnt;
    q.mnb = xparams.mnb;
    q.n_lo = nlo_b;
    q.n_hi = nhi_b;
    q.n_mix = nmix_b;
    q.n_dmix = xparams.n_dmix;
    q.ndmix_tab = xparams.ndmix_tab;
    q.nlo = xparams.nlo;
    q.L = L;
    q.pair_mhi = pair;
    q.lds_x_off = tab_x;
    xp_lds = int(lds);
    if (xp_lds > 64 * 1024)
        QK_HIP(hipFuncSetAttribute((const void*)k_xrow, hipFuncAttributeMaxDynamicSharedMemorySize, xp_lds));
    xp_groups_n = u32(groups.size());
    xp_ctiles_n = u32(cbands.size());
    if (xp_lds_c > 64 * 1024)
        QK_HIP(hipFuncSetAttribute((const void*)k_xcol, hipFuncAttributeMaxDynamicSharedMemorySize, xp_lds_c));
    xpass_ok = true;
    // off until autotune() finds it faster (QKRYLOV_XPASS=1 / =0 forces it)
    xpass = false;
    if (const char* e = std::getenv("QKRYLOV_XPASS")) xpass = std::atoi(e) != 0;
}

void Hamiltonian::apply_xpass(const cplx* x, cplx* y, double* dot_part, hipStream_t st) const {
    XPassParams q = xpp;
    hipLaunchKernelGGL(k_xcol, dim3(xp_grid_c), dim3(kXpThreads), size_t(xp_lds_c), st, q, x, y);
    q.dot_part = dot_part;
    hipLaunchKernelGGL(k_xrow, dim3(xp_grid_r), dim3(kXpThreads), size_t(xp_lds), st, q, x, y);
    QK_HIP(hipGetLastError());
}

}  // namespace qk
