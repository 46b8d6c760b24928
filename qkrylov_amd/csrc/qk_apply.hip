// y = H x, matrix-free, on gfx950.
//
// Work decomposition.  The basis is a ragged 2-D array: row = high half `hi`
// of the state word, column = rank r of the low half inside its class
// (see Basis in qk_model.hpp).  A tile is 64 consecutive columns of one row
// and is processed by one wave64, so `hi`, the row offset jhi[hi] and every
// high-half quantity are wave-uniform (SGPRs, scalar loads).  Lanes hold one
// basis state each: lo = lo_list[class][r], index a = jhi[hi] + r.
//
// For every compiled group (qk_model.hpp: flip mask F, Jordan-Wigner mask J,
// support bits, conj-amplitude table) a lane reads its configuration on the
// support, looks the amplitude up in LDS and, if non-zero, ranks the target
// t = s ^ F:
//   diagonal  F == 0          : contributes amp * x[a]
//   low       F inside lo     : index = jhi[hi] + jlo[lo ^ F]     (jlo in LDS)
//   high      F inside hi     : index = jhi[hi ^ F>>L] + r        (uniform row,
//                                                                  coalesced)
//   mixed     F on both halves: index = jhi[hi'] + jlo[lo']
// and the target must be in the basis (class of lo' == class required by hi').
// y is written once per state (gather form, no atomics, no zero-fill), the
// reference's convention y[a] = sum conj(amp(a->b)) x[b]
// (src/hamiltonian/matrix_free_hamiltonian.cpp:75-84, 156-167).
//
// Three kernels share this decomposition: the exchange kernels (every
// off-diagonal group a real two-bit exchange: bonds instead of groups,
// diagonals from per-half tables), the row-program kernel (k_apply with
// ROWS: groups inside one half precomputed per class or per row, groups
// across the split as per-row entries over low-half sub-tables) and the
// group loop above for everything else.  All three add a state's terms in
// group order, so block and single-vector launches give identical bits.
//
// Tiles are dealt to the 8 XCDs in 8 contiguous chunks (workgroup i runs on
// XCD i % 8), so each XCD's L2 holds the rows its waves are reading.
#include "qk_hamiltonian.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <cstdlib>
#include <sstream>
#include <string>

namespace qk {

namespace {

constexpr int kWaves = 4;            // waves per workgroup (diagonal, S^z(q) kernels)
constexpr int kBlock = 64 * kWaves;  // threads per workgroup
constexpr int kGWaves = 16;          // waves per workgroup (generic apply): one LDS
                                     // copy of the tables serves 16 waves
constexpr int kGBlock = 64 * kGWaves;
constexpr int kXWaves = 16;          // waves per workgroup (exchange kernel)
constexpr int kXcd = 8;              // XCDs on MI355X

__device__ __forceinline__ void cfma(cplx& acc, const cplx a, const cplx x) {
    acc.re += a.re * x.re - a.im * x.im;
    acc.im += a.re * x.im + a.im * x.re;
}

template <typename S>
__device__ __forceinline__ int popc_s(S v) {
    if constexpr (sizeof(S) == 8) return __popcll(v);
    else return __popc(v);
}

template <typename S>
__device__ __forceinline__ u32 config_of(S s, const DevGroup& g) {
    u32 c = 0;
    u64 pw = g.posw;
    for (u32 k = 0; k < g.nsup; ++k, pw >>= 6) c |= u32((s >> u32(pw & 63)) & 1) << k;
    return c;
}

// Tile schedule.  The tile list is cut into chunks of `chunk` tiles; chunk c
// belongs to XCD c % 8 (workgroup b runs on XCD b % 8).  The waves of one
// XCD sweep its chunks in list order, so at any moment the 8 XCDs work on 8
// neighbouring chunks of the (class-major) list: each XCD's L2 holds the
// rows of its chunk and the MALL holds the current class.
struct TileIter {
    u32 next_id, nchunks, w, W, chunk, ntiles, t, t_end;
    __device__ __forceinline__ bool next_chunk() {
        while (next_id < nchunks) {
            const u32 c = next_id;
            next_id += kXcd;
            t = c * chunk + w;
            t_end = min((c + 1) * chunk, ntiles);
            if (t < t_end) return true;
        }
        return false;
    }
};

template <int WV = kWaves, typename PP>
__device__ __forceinline__ TileIter tile_iter(const PP& p, u32 wave) {
    TileIter it;
    it.w = (blockIdx.x / kXcd) * WV + wave;
    it.W = (gridDim.x / kXcd) * WV;
    it.chunk = p.tiles_per_xcd;   // tiles per chunk
    it.ntiles = p.ntiles;
    it.next_id = blockIdx.x % kXcd;
    it.nchunks = (p.ntiles + it.chunk - 1) / it.chunk;
    it.t = it.t_end = 0;
    return it;
}

// Iterate the tiles of this wave: QK_FOR_TILES(p, wave, t) { ... }
#define QK_FOR_TILES(p, wave, t)                                             \
    for (TileIter it_ = tile_iter(p, wave); it_.next_chunk();)              \
        for (u32 t = it_.t; t < it_.t_end; t += it_.W)
#define QK_FOR_TILES_W(p, wave, t, WV)                                       \
    for (TileIter it_ = tile_iter<WV>(p, wave); it_.next_chunk();)          \
        for (u32 t = it_.t; t < it_.t_end; t += it_.W)

template <typename T>
__device__ __forceinline__ void store_nt(T* p, const T& v) {
    __builtin_nontemporal_store(v.re, &p->re);
    __builtin_nontemporal_store(v.im, &p->im);
}

// CROSS: rows (y) enumerate one basis and the targets (x) are ranked in
// another of the same kind, size and counted masks (ColTables: its row
// offsets and row classes; the low-half ranks are shared).  y has stride
// c.ld_out, x stride ld.
// ROWS: the groups inside one half come from the row-program tables (q) and
// p holds only the others; contributions are added in group order within
// each kind.
template <typename S, int B, bool CROSS, bool ROWS>
__global__ void __launch_bounds__(kGBlock)
k_apply(const ApplyParams p, const ColTables c, const RowParams q, const cplx* __restrict__ x,
        cplx* __restrict__ y, const u64 ld) {
    static_assert(!(CROSS && ROWS), "row programs are built for one basis");
    extern __shared__ __align__(16) unsigned char smem[];
    cplx* s_tab = reinterpret_cast<cplx*>(smem);
    u16* s_jlo = reinterpret_cast<u16*>(smem + size_t(p.ntab) * sizeof(cplx));
    cplx* s_amp = reinterpret_cast<cplx*>(smem + size_t(p.ntab) * sizeof(cplx) +
                                          ((size_t(p.nlo) * sizeof(u16) + 15) / 16) * 16);
    DevGroup* s_mdesc = reinterpret_cast<DevGroup*>(s_amp + (ROWS ? q.namp : 0u));

    for (u32 i = threadIdx.x; i < p.ntab; i += kGBlock) s_tab[i] = p.tables[i];
    if constexpr (ROWS) {
        for (u32 i = threadIdx.x; i < q.namp; i += kGBlock) s_amp[i] = q.amps[i];
        for (u32 i = threadIdx.x; i < q.nmdesc; i += kGBlock) s_mdesc[i] = q.mdesc[i];
    }
    if ((p.nlo & 7) == 0) {
        const uint4* src = reinterpret_cast<const uint4*>(p.jlo);
        uint4* dst = reinterpret_cast<uint4*>(s_jlo);
        for (u32 i = threadIdx.x; i < p.nlo / 8; i += kGBlock) dst[i] = src[i];
    } else {
        for (u32 i = threadIdx.x; i < p.nlo; i += kGBlock) s_jlo[i] = p.jlo[i];
    }
    __syncthreads();

    const u32 wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const u32 lane = threadIdx.x & 63;
    const S lomask = (S(1) << p.L) - 1;

    QK_FOR_TILES_W(p, wave, t, kGWaves) {
        const uint2 td = p.tiles[t];
        const u32 hi = __builtin_amdgcn_readfirstlane(td.x);
        const u32 r0 = __builtin_amdgcn_readfirstlane(td.y);
        const int cls = p.hi_cls[hi];
        const u64 base = p.jhi[hi];
        const u32 cnt = p.cls_cnt[cls];
        const u32 off = p.cls_off[cls];
        const u32 r = r0 + lane;
        if (r >= cnt) continue;
        const u32 lo = p.lo_list[off + r];
        const S s = (S(hi) << p.L) | S(lo);
        const u64 a = base + r;
        // row hi of the target basis: same-row targets and the diagonal
        u64 base_c = base;
        int cls_c = cls;
        if constexpr (CROSS) {
            base_c = c.jhi[hi];
            cls_c = base_c == ~u64(0) ? -1 : c.hi_cls[hi];
        }

        cplx acc[B];
#pragma unroll
        for (int v = 0; v < B; ++v) acc[v] = cplx{0.0, 0.0};
        cplx dsum{0.0, 0.0};
        // row-program records: the first batch of the row's high-half and
        // across-the-split entries is loaded here, so that its latency
        // overlaps the low-half work; each later batch is loaded while the
        // gathers of the one before are in flight
        constexpr int NE = B >= 4 ? 2 : 4;
        u32 vr = 0, he0 = 0, he1 = 0, me0 = 0, me1 = 0;
        RowEntry R[NE];
        RowMix MR[NE];
        if constexpr (ROWS) {
            vr = q.vrow ? q.vrow[hi] : hi;
            const cplx d0 = q.dlo[off + r], d1 = q.dhi[vr];
            dsum.re = d0.re + d1.re;
            dsum.im = d0.im + d1.im;
            he0 = q.rptr[vr];
            he1 = q.rptr[vr + 1];
#pragma unroll
            for (int k = 0; k < NE; ++k)
                R[k] = he0 + k < he1 ? q.rent[he0 + k] : RowEntry{0, cplx{0.0, 0.0}};
            if (q.ment) {
                me0 = q.mptr[vr];
                me1 = q.mptr[vr + 1];
#pragma unroll
                for (int k = 0; k < NE; ++k) MR[k] = me0 + k < me1 ? q.ment[me0 + k] : RowMix{};
            }
        }

        u32 g = 0;
        // diagonal groups
        for (; g < p.n_diag; ++g) {
            const DevGroup G = p.groups[g];
            cplx amp = s_tab[G.tab + config_of(s, G)];
            if (G.jw && (popc_s(s & S(G.jw)) & 1)) { amp.re = -amp.re; amp.im = -amp.im; }
            dsum.re += amp.re;
            dsum.im += amp.im;
        }
        if constexpr (ROWS) {
            // low-half groups of the row program: this rank's entries
            // entries are read C at a time ahead of their gathers; a lane's
            // padding entries (after its last one) end its loop
            constexpr int C = B >= 8 ? 1 : (B >= 4 ? 2 : 4);
            const u32 w = q.ell_w[cls];
            const u32* __restrict__ E = q.ell + q.ell_off[cls] + r;
            for (u32 k = 0; k < w; k += C) {
                u32 e[C];
#pragma unroll
                for (int j = 0; j < C; ++j) e[j] = (k + j < w) ? E[size_t(k + j) * cnt] : ~0u;
                if (e[0] == ~0u) break;
#pragma unroll
                for (int j = 0; j < C; ++j) {
                    if (e[j] == ~0u) break;
                    const cplx amp = s_amp[e[j] >> 16];
                    const u64 b = base + (e[j] & 0xFFFFu);
#pragma unroll
                    for (int v = 0; v < B; ++v) cfma(acc[v], amp, x[v * ld + b]);
                }
            }
        }
        // low-half groups: same row
        const u32 g_lo_end = p.n_diag + p.n_lo;
        for (; g < g_lo_end; ++g) {
            const DevGroup G = p.groups[g];
            cplx amp = s_tab[G.tab + config_of(s, G)];
            if (amp.re == 0.0 && amp.im == 0.0) continue;
            const u32 tlo = lo ^ u32(G.flip);
            const u16 rl = s_jlo[tlo];
            if (rl == 0xFFFF) continue;
            if (__popc(tlo & u32(p.m1lo)) * p.ncls2 + __popc(tlo & u32(p.m2lo)) != cls_c) continue;
            if (G.jw && (popc_s(s & S(G.jw)) & 1)) { amp.re = -amp.re; amp.im = -amp.im; }
            const u64 b = base_c + rl;
#pragma unroll
            for (int v = 0; v < B; ++v) cfma(acc[v], amp, x[v * ld + b]);
        }
        if constexpr (ROWS) {
            // high-half groups of the row program: the row's entries
            // (uniform), NE at a time with all their gathers issued before
            // the first is used; added in entry order.  A slot past the
            // row's last entry reads its own state and adds nothing.
            for (u32 e = he0; e < he1; e += NE) {
                cplx xv[NE][B];
                cplx amp[NE];
#pragma unroll
                for (int k = 0; k < NE; ++k) {
                    const bool on = e + k < he1;
                    amp[k] = R[k].amp;
                    const u64 b = on ? R[k].base + r : a;
#pragma unroll
                    for (int v = 0; v < B; ++v) xv[k][v] = x[v * ld + b];
                }
#pragma unroll
                for (int k = 0; k < NE; ++k)
                    R[k] = e + NE + k < he1 ? q.rent[e + NE + k] : RowEntry{0, cplx{0.0, 0.0}};
#pragma unroll
                for (int k = 0; k < NE; ++k) {
                    if (e + k >= he1) break;
#pragma unroll
                    for (int v = 0; v < B; ++v) cfma(acc[v], amp[k], xv[k][v]);
                }
            }
        }
        // high-half groups: another row, same column
        const u32 g_hi_end = g_lo_end + p.n_hi;
        for (; g < g_hi_end; ++g) {
            const DevGroup G = p.groups[g];
            const u32 hi2 = hi ^ u32(G.flip >> p.L);
            const u64 base2 = CROSS ? c.jhi[hi2] : p.jhi[hi2];
            if (base2 == ~u64(0) || (CROSS ? c.hi_cls[hi2] : p.hi_cls[hi2]) != cls) continue;   // uniform
            cplx amp = s_tab[G.tab + config_of(s, G)];
            if (amp.re == 0.0 && amp.im == 0.0) continue;
            if (G.jw && (popc_s(s & S(G.jw)) & 1)) { amp.re = -amp.re; amp.im = -amp.im; }
            const u64 b = base2 + r;
#pragma unroll
            for (int v = 0; v < B; ++v) cfma(acc[v], amp, x[v * ld + b]);
        }
        // mixed groups
        const u32 g_end = g_hi_end + p.n_mix;
        for (; g < g_end; ++g) {
            const DevGroup G = p.groups[g];
            const u32 hi2 = hi ^ u32(G.flip >> p.L);
            const u64 base2 = CROSS ? c.jhi[hi2] : p.jhi[hi2];
            if (base2 == ~u64(0)) continue;                               // uniform
            const int cls2 = CROSS ? c.hi_cls[hi2] : p.hi_cls[hi2];
            cplx amp = s_tab[G.tab + config_of(s, G)];
            if (amp.re == 0.0 && amp.im == 0.0) continue;
            const u32 tlo = lo ^ u32(G.flip & u64(lomask));
            const u16 rl = s_jlo[tlo];
            if (rl == 0xFFFF) continue;
            if (__popc(tlo & u32(p.m1lo)) * p.ncls2 + __popc(tlo & u32(p.m2lo)) != cls2) continue;
            if (G.jw && (popc_s(s & S(G.jw)) & 1)) { amp.re = -amp.re; amp.im = -amp.im; }
            const u64 b = base2 + rl;
#pragma unroll
            for (int v = 0; v < B; ++v) cfma(acc[v], amp, x[v * ld + b]);
        }
        if constexpr (ROWS) {
            if (q.ment) {
                // groups across the split: the row's entries (uniform), each
                // finished per lane on the low half; NE entries at a time
                // (records, then every lane's gathers, then the sums, in
                // entry order).  A lane whose target is not in the basis
                // reads its own state and adds zero.
                for (u32 e = me0; e < me1; e += NE) {
                    cplx amp[NE];
                    u64 tb[NE];
#pragma unroll
                    for (int k = 0; k < NE; ++k) {
                        amp[k] = cplx{0.0, 0.0};
                        tb[k] = a;
                        if (e + k >= me1) continue;
                        const RowMix M = MR[k];
                        const DevGroup D = s_mdesc[M.desc_sign >> 1];
                        cplx am = s_tab[M.tab + config_of(lo, D)];
                        const u32 tlo = lo ^ u32(D.flip);
                        const u16 rl = s_jlo[tlo];
                        const bool ok = (am.re != 0.0 || am.im != 0.0) && rl != 0xFFFF &&
                                        __popc(tlo & u32(p.m1lo)) * p.ncls2 + __popc(tlo & u32(p.m2lo)) == M.cls;
                        if (!ok) continue;
                        if ((__popc(lo & u32(D.jw)) ^ M.desc_sign) & 1) { am.re = -am.re; am.im = -am.im; }
                        amp[k] = am;
                        tb[k] = M.base + rl;
                    }
                    cplx xv[NE][B];
#pragma unroll
                    for (int k = 0; k < NE; ++k)
#pragma unroll
                        for (int v = 0; v < B; ++v) xv[k][v] = x[v * ld + tb[k]];
#pragma unroll
                    for (int k = 0; k < NE; ++k) MR[k] = e + NE + k < me1 ? q.ment[e + NE + k] : RowMix{};
#pragma unroll
                    for (int k = 0; k < NE; ++k) {
                        if (amp[k].re == 0.0 && amp[k].im == 0.0) continue;
#pragma unroll
                        for (int v = 0; v < B; ++v) cfma(acc[v], amp[k], xv[k][v]);
                    }
                }
            }
        }
#pragma unroll
        for (int v = 0; v < B; ++v) {
            cplx o = acc[v];
            if ((p.n_diag || (ROWS && q.ndiag)) && cls_c == cls) cfma(o, dsum, x[v * ld + base_c + r]);
            if constexpr (CROSS) y[v * c.ld_out + a] = o;
            else y[v * ld + a] = o;
        }
    }
}

// Exchange fast path (see XBond).  Per lane: the diagonal from two table
// loads (low-half terms, aligned with lo_list; high-half terms per row,
// uniform) plus the few terms that straddle the split; then each bond with
// differing bits contributes a*x[target].  Low bonds rank the target through
// the LDS table, high bonds are wave-uniform (one scalar test and one scalar
// row lookup, then a coalesced 64-lane read of the other row), mixed bonds do
// both.  Exchange conserves the counted popcount, so no class test is needed.
// Workgroups of kXWaves waves share one LDS copy of the rank table, so the
// 32 KB table costs LDS once per 16 waves and the CU holds 8 waves per SIMD.
template <int B>
__global__ void __launch_bounds__(64 * kXWaves)
k_apply_xchg(const XParams p, const cplx* __restrict__ x, cplx* __restrict__ y, const u64 ld) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr u32 NT = 64 * kXWaves;
    cplx* s_tab = reinterpret_cast<cplx*>(smem);
    u16* s_jlo = reinterpret_cast<u16*>(smem + size_t(p.ndmix_tab) * sizeof(cplx));
    for (u32 i = threadIdx.x; i < p.ndmix_tab; i += NT) s_tab[i] = p.dmix_tab[i];
    if ((p.nlo & 7) == 0) {
        const uint4* src = reinterpret_cast<const uint4*>(p.jlo);
        uint4* dst = reinterpret_cast<uint4*>(s_jlo);
        for (u32 i = threadIdx.x; i < p.nlo / 8; i += NT) dst[i] = src[i];
    } else {
        for (u32 i = threadIdx.x; i < p.nlo; i += NT) s_jlo[i] = p.jlo[i];
    }
    __syncthreads();

    const u32 wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const u32 lane = threadIdx.x & 63;
    const u32 hi_end = p.n_lo + p.n_hi;
    const u32 mix_end = hi_end + p.n_mix;

    QK_FOR_TILES_W(p, wave, t, kXWaves) {
        const XTile T = p.xtiles[t];                  // one scalar load
        const u32 hi = __builtin_amdgcn_readfirstlane(T.hi);
        const u64 base = T.base;
        if (lane >= T.nvalid) continue;
        const u32 r = T.r0 + lane;
        const u32 lo = p.lo_list[T.pos0 + lane];
        const u64 a = base + r;

        double dg = p.dlo[T.pos0 + lane] + T.dh;
        if (p.n_dmix) {
            const u64 s = (u64(hi) << p.L) | lo;
            for (u32 g = 0; g < p.n_dmix; ++g) {
                const DevGroup G = p.dmix[g];
                dg += s_tab[G.tab + config_of(s, G)].re;
            }
        }
        cplx acc[B];
#pragma unroll
        for (int v = 0; v < B; ++v) {
            const cplx xa = x[v * ld + a];
            acc[v] = cplx{dg * xa.re, dg * xa.im};
        }
        u32 b = 0;
        for (; b < p.n_lo; ++b) {
            const XBond X = p.bonds[b];
            const u32 vi = (lo >> X.bi) & 1, vj = (lo >> X.bj) & 1;
            if (vi == vj) continue;
            const double amp = vi ? X.a10 : X.a01;
            const u64 idx = base + s_jlo[lo ^ X.mlo];
#pragma unroll
            for (int v = 0; v < B; ++v) {
                const cplx xv = x[v * ld + idx];
                acc[v].re += amp * xv.re;
                acc[v].im += amp * xv.im;
            }
        }
        for (; b < hi_end; ++b) {
            const XBond X = p.bonds[b];
            const u32 vi = (hi >> (X.bi - p.L)) & 1, vj = (hi >> (X.bj - p.L)) & 1;
            if (vi == vj) continue;                          // uniform
            const double amp = vi ? X.a10 : X.a01;
            const u64 idx = p.jhi[hi ^ X.mhi] + r;
#pragma unroll
            for (int v = 0; v < B; ++v) {
                const cplx xv = x[v * ld + idx];
                acc[v].re += amp * xv.re;
                acc[v].im += amp * xv.im;
            }
        }
        for (; b < mix_end; ++b) {
            const XBond X = p.bonds[b];
            const u32 vj = (hi >> (X.bj - p.L)) & 1;       // uniform
            const u64 base2 = p.jhi[hi ^ X.mhi];            // uniform
            const u32 vi = (lo >> X.bi) & 1;
            if (vi == vj || base2 == ~u64(0)) continue;
            const double amp = vi ? X.a10 : X.a01;
            const u64 idx = base2 + s_jlo[lo ^ X.mlo];
#pragma unroll
            for (int v = 0; v < B; ++v) {
                const cplx xv = x[v * ld + idx];
                acc[v].re += amp * xv.re;
                acc[v].im += amp * xv.im;
            }
        }
#pragma unroll
        for (int v = 0; v < B; ++v) store_nt(&y[v * ld + a], acc[v]);
    }
}

// Multi-sector exchange kernel: the tiles of several sectors of one model in
// one launch, each tile applied to all vectors of its sector, B at a time.
// Ranks, bond tests and diagonal of a state are computed once per tile and
// reused for every vector; only the row offsets (jhi) differ by sector.
template <int B>
__global__ void __launch_bounds__(64 * kXWaves)
k_apply_xchg_multi(const XParams p, const MParams mp, const cplx* __restrict__ x,
                   cplx* __restrict__ y) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr u32 NT = 64 * kXWaves;
    cplx* s_tab = reinterpret_cast<cplx*>(smem);
    u16* s_jlo = reinterpret_cast<u16*>(smem + size_t(p.ndmix_tab) * sizeof(cplx));
    for (u32 i = threadIdx.x; i < p.ndmix_tab; i += NT) s_tab[i] = p.dmix_tab[i];
    if ((p.nlo & 7) == 0) {
        const uint4* src = reinterpret_cast<const uint4*>(p.jlo);
        uint4* dst = reinterpret_cast<uint4*>(s_jlo);
        for (u32 i = threadIdx.x; i < p.nlo / 8; i += NT) dst[i] = src[i];
    } else {
        for (u32 i = threadIdx.x; i < p.nlo; i += NT) s_jlo[i] = p.jlo[i];
    }
    __syncthreads();

    const u32 wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const u32 lane = threadIdx.x & 63;
    const u32 hi_end = p.n_lo + p.n_hi;
    const u32 mix_end = hi_end + p.n_mix;

    QK_FOR_TILES_W(mp, wave, t, kXWaves) {
        const MTile T = mp.tiles[t];
        const u32 hi = __builtin_amdgcn_readfirstlane(T.hi);
        const u64* __restrict__ jhi = mp.jhis[T.sec];
        if (lane >= T.nvalid) continue;
        const u32 r = T.r0 + lane;
        const u32 lo = p.lo_list[T.pos0 + lane];
        const u64 a = T.base + r;
        double dg = p.dlo[T.pos0 + lane] + T.dh;
        if (p.n_dmix) {
            const u64 s = (u64(hi) << p.L) | lo;
            for (u32 g = 0; g < p.n_dmix; ++g) {
                const DevGroup G = p.dmix[g];
                dg += s_tab[G.tab + config_of(s, G)].re;
            }
        }
        for (u32 v0 = 0; v0 < T.nvec; v0 += B) {
            const u32 nv = min(u32(B), T.nvec - v0);
            const cplx* xs = x + T.xoff + u64(v0) * T.ld;
            cplx* ys = y + T.xoff + u64(v0) * T.ld;
            cplx acc[B];
#pragma unroll
            for (int v = 0; v < B; ++v) {
                acc[v] = cplx{0.0, 0.0};
                if (u32(v) < nv) {
                    const cplx xa = xs[v * T.ld + a];
                    acc[v] = cplx{dg * xa.re, dg * xa.im};
                }
            }
            u32 b = 0;
            for (; b < p.n_lo; ++b) {
                const XBond X = p.bonds[b];
                const u32 vi = (lo >> X.bi) & 1, vj = (lo >> X.bj) & 1;
                if (vi == vj) continue;
                const double amp = vi ? X.a10 : X.a01;
                const u64 idx = T.base + s_jlo[lo ^ X.mlo];
#pragma unroll
                for (int v = 0; v < B; ++v)
                    if (u32(v) < nv) {
                        const cplx xv = xs[v * T.ld + idx];
                        acc[v].re += amp * xv.re;
                        acc[v].im += amp * xv.im;
                    }
            }
            for (; b < hi_end; ++b) {
                const XBond X = p.bonds[b];
                const u32 vi = (hi >> (X.bi - p.L)) & 1, vj = (hi >> (X.bj - p.L)) & 1;
                if (vi == vj) continue;                      // uniform
                const double amp = vi ? X.a10 : X.a01;
                const u64 idx = jhi[hi ^ X.mhi] + r;
#pragma unroll
                for (int v = 0; v < B; ++v)
                    if (u32(v) < nv) {
                        const cplx xv = xs[v * T.ld + idx];
                        acc[v].re += amp * xv.re;
                        acc[v].im += amp * xv.im;
                    }
            }
            for (; b < mix_end; ++b) {
                const XBond X = p.bonds[b];
                const u32 vj = (hi >> (X.bj - p.L)) & 1;   // uniform
                const u64 base2 = jhi[hi ^ X.mhi];          // uniform
                const u32 vi = (lo >> X.bi) & 1;
                if (vi == vj || base2 == ~u64(0)) continue;
                const double amp = vi ? X.a10 : X.a01;
                const u64 idx = base2 + s_jlo[lo ^ X.mlo];
#pragma unroll
                for (int v = 0; v < B; ++v)
                    if (u32(v) < nv) {
                        const cplx xv = xs[v * T.ld + idx];
                        acc[v].re += amp * xv.re;
                        acc[v].im += amp * xv.im;
                    }
            }
#pragma unroll
            for (int v = 0; v < B; ++v)
                if (u32(v) < nv) store_nt(&ys[v * T.ld + a], acc[v]);
        }
    }
}

// Bond batches of the single-vector exchange kernel.  Every load of a batch
// (bond records, LDS ranks, the SPL gathers of all NB bonds) is issued before
// the first one is used, and a lane without a contribution loads nothing and
// adds zero, so a wave keeps NB*SPL gathers in flight instead of waiting out
// one bond at a time.  Contributions are added in bond order, as in the
// one-bond loop.  SYM: every bond has one amplitude for both directions
// (a01 == a10, e.g. real exchange couplings), taken straight from the bond
// record instead of selected per lane; a lane without a contribution adds
// a * 0.

// low-half bonds: same row, rank from the LDS table
template <int SPL, int NB, bool SYM>
__device__ __forceinline__ void xbonds_lo(const XBond* __restrict__ bonds, const u16* s_jlo,
                                          const cplx* __restrict__ x, const u64 base,
                                          const u32 (&lo)[SPL], const bool (&v)[SPL],
                                          cplx (&acc)[SPL]) {
    XBond X[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) X[u] = bonds[u];
    double amp[NB][SPL];
    u32 idx[NB][SPL];
    bool ok[NB][SPL];
#pragma unroll
    for (int u = 0; u < NB; ++u)
#pragma unroll
        for (int k = 0; k < SPL; ++k) {
            const u32 vi = (lo[k] >> X[u].bi) & 1, vj = (lo[k] >> X[u].bj) & 1;
            ok[u][k] = v[k] && vi != vj;
            if constexpr (SYM) amp[u][k] = X[u].a01;
            else amp[u][k] = ok[u][k] ? (vi ? X[u].a10 : X[u].a01) : 0.0;
            const u32 t = s_jlo[lo[k] ^ X[u].mlo];   // lo ^ mlo < 2^L: in the table
            idx[u][k] = SYM ? t : (ok[u][k] ? t : 0u);
        }
    cplx xv[NB][SPL];
#pragma unroll
    for (int u = 0; u < NB; ++u)
#pragma unroll
        for (int k = 0; k < SPL; ++k) {
            xv[u][k] = cplx{0.0, 0.0};
            if (ok[u][k]) xv[u][k] = x[base + idx[u][k]];
        }
#pragma unroll
    for (int u = 0; u < NB; ++u)
#pragma unroll
        for (int k = 0; k < SPL; ++k) {
            acc[k].re += amp[u][k] * xv[u][k].re;
            acc[k].im += amp[u][k] * xv[u][k].im;
        }
}

// high-half bonds: the row's active bonds (XRowNb, wave-uniform), same ranks
// in the partner row.  A lane past the tile's end reads its wave's first rank.
template <int SPL, int NB>
__device__ __forceinline__ void xbonds_hi(const XRowNb* __restrict__ nb, const cplx* __restrict__ x,
                                          const u32 (&r)[SPL], const bool (&v)[SPL],
                                          cplx (&acc)[SPL]) {
    XRowNb E[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) E[u] = nb[u];
    cplx xv[NB][SPL];
#pragma unroll
    for (int u = 0; u < NB; ++u)
#pragma unroll
        for (int k = 0; k < SPL; ++k) xv[u][k] = x[E[u].base + (v[k] ? r[k] : r[0])];
#pragma unroll
    for (int u = 0; u < NB; ++u)
#pragma unroll
        for (int k = 0; k < SPL; ++k) {
            acc[k].re += E[u].amp * xv[u][k].re;
            acc[k].im += E[u].amp * xv[u][k].im;
        }
}

// bonds straddling the split: another row (offset from the per-row record)
// and another low-half class
template <int SPL, int NB, bool SYM>
__device__ __forceinline__ void xbonds_mix(const XBond* __restrict__ bonds, const u64* __restrict__ mnb,
                                           const u16* s_jlo, const u32 hi, const u32 L,
                                           const cplx* __restrict__ x, const u64 base,
                                           const u32 (&lo)[SPL], const bool (&v)[SPL],
                                           cplx (&acc)[SPL]) {
    XBond X[NB];
    u64 b2[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        X[u] = bonds[u];
        b2[u] = mnb[u];
    }
    double amp[NB][SPL];
    u64 idx[NB][SPL];
    bool ok[NB][SPL];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        const u32 vj = (hi >> (X[u].bj - L)) & 1;
#pragma unroll
        for (int k = 0; k < SPL; ++k) {
            const u32 vi = (lo[k] >> X[u].bi) & 1;
            ok[u][k] = v[k] && b2[u] != ~u64(0) && vi != vj;
            if constexpr (SYM) amp[u][k] = X[u].a01;
            else amp[u][k] = ok[u][k] ? (vi ? X[u].a10 : X[u].a01) : 0.0;
            const u32 t = s_jlo[lo[k] ^ X[u].mlo];
            idx[u][k] = SYM ? b2[u] + t : (ok[u][k] ? b2[u] + t : base);
        }
    }
    cplx xv[NB][SPL];
#pragma unroll
    for (int u = 0; u < NB; ++u)
#pragma unroll
        for (int k = 0; k < SPL; ++k) {
            xv[u][k] = cplx{0.0, 0.0};
            if (ok[u][k]) xv[u][k] = x[idx[u][k]];
        }
#pragma unroll
    for (int u = 0; u < NB; ++u)
#pragma unroll
        for (int k = 0; k < SPL; ++k) {
            acc[k].re += amp[u][k] * xv[u][k].re;
            acc[k].im += amp[u][k] * xv[u][k].im;
        }
}

// LDS of the exchange kernels: the diagonal tables of the groups spanning
// both halves, then the low-half rank table.
template <int WV = kXWaves>
__device__ __forceinline__ void xstage_lds(const XParams& p, unsigned char* smem, cplx*& s_tab,
                                           u16*& s_jlo) {
    constexpr u32 NT = 64 * WV;
    s_tab = reinterpret_cast<cplx*>(smem);
    s_jlo = reinterpret_cast<u16*>(smem + size_t(p.ndmix_tab) * sizeof(cplx));
    for (u32 i = threadIdx.x; i < p.ndmix_tab; i += NT) s_tab[i] = p.dmix_tab[i];
    if ((p.nlo & 7) == 0) {
        const uint4* src = reinterpret_cast<const uint4*>(p.jlo);
        uint4* dst = reinterpret_cast<uint4*>(s_jlo);
        for (u32 i = threadIdx.x; i < p.nlo / 8; i += NT) dst[i] = src[i];
    } else {
        for (u32 i = threadIdx.x; i < p.nlo; i += NT) s_jlo[i] = p.jlo[i];
    }
    __syncthreads();
}

// One tile of the single-vector exchange kernels (lane < T.nvalid): y over
// the tile's states, x and y at the vector's start, bond records R of the
// vector's sector.  Adds Re<x, y> over this lane's states to dot.
template <int SPL, int NB, bool SYM>
__device__ __forceinline__ void xs_tile(const XParams& p, const XTile& T, const u32 hi,
                                        const XRecs& R, const cplx* __restrict__ x,
                                        cplx* __restrict__ y, const cplx* s_tab,
                                        const u16* s_jlo, const u32 lane, double& dot) {
    const u32 hi_end = p.n_lo + p.n_hi;
    const u32 mix_end = hi_end + p.n_mix;
    const u64 base = T.base;
    __builtin_assume(lane < T.nvalid);   // the caller skips lanes past the tile
    u32 lo[SPL], r[SPL];
    bool v[SPL];
    cplx acc[SPL], xa[SPL];
#pragma unroll
    for (int k = 0; k < SPL; ++k) {
        const u32 j = lane + 64u * k;
        v[k] = j < T.nvalid;
        r[k] = T.r0 + j;
        lo[k] = v[k] ? p.lo_list[T.pos0 + j] : 0u;
        double dg = v[k] ? p.dlo[T.pos0 + j] + T.dh : 0.0;
        if (p.n_dmix) {
            const u64 s = (u64(hi) << p.L) | lo[k];
            for (u32 g = 0; g < p.n_dmix; ++g) {
                const DevGroup G = p.dmix[g];
                dg += s_tab[G.tab + config_of(s, G)].re;
            }
        }
        xa[k] = cplx{0.0, 0.0};
        if (v[k]) xa[k] = x[base + r[k]];
        acc[k] = cplx{dg * xa[k].re, dg * xa[k].im};
    }
    u32 b = 0;
    for (; b + NB <= p.n_lo; b += NB) xbonds_lo<SPL, NB, SYM>(p.bonds + b, s_jlo, x, base, lo, v, acc);
    for (; b < p.n_lo; ++b) xbonds_lo<SPL, 1, SYM>(p.bonds + b, s_jlo, x, base, lo, v, acc);
    const u32 vr = R.vrow ? R.vrow[hi] : hi;   // record slot of this row
    {
        const u32 hc = R.hcnt[vr];
        const XRowNb* nb = R.hnb + size_t(vr) * p.n_hi;
        u32 e = 0;
        for (; e + NB <= hc; e += NB) xbonds_hi<SPL, NB>(nb + e, x, r, v, acc);
        for (; e < hc; ++e) xbonds_hi<SPL, 1>(nb + e, x, r, v, acc);
    }
    b = hi_end;
    const u64* mnb = R.mnb + size_t(vr) * p.n_mix;
    for (; b + NB <= mix_end; b += NB)
        xbonds_mix<SPL, NB, SYM>(p.bonds + b, mnb + (b - hi_end), s_jlo, hi, p.L, x, base, lo, v, acc);
    for (; b < mix_end; ++b)
        xbonds_mix<SPL, 1, SYM>(p.bonds + b, mnb + (b - hi_end), s_jlo, hi, p.L, x, base, lo, v, acc);
#pragma unroll
    for (int k = 0; k < SPL; ++k)
        if (v[k]) {
            store_nt(&y[base + r[k]], acc[k]);
            dot += xa[k].re * acc[k].re + xa[k].im * acc[k].im;
        }
}

// Single-vector exchange kernel with SPL states per lane: a tile spans
// 64*SPL consecutive ranks of one row and lane l holds ranks r0+l+64k.
// Bonds are taken NB at a time (xbonds_*), so up to NB*SPL gathers per wave
// are in flight.  SYM: every low-half and straddling bond has a01 == a10.
// WV waves per workgroup.
template <int SPL, int NB, bool SYM, int WV = kXWaves>
__global__ void __launch_bounds__(64 * WV)
k_apply_xchg_s(const XParams p, const cplx* __restrict__ x, cplx* __restrict__ y) {
    extern __shared__ __align__(16) unsigned char smem[];
    cplx* s_tab;
    u16* s_jlo;
    xstage_lds<WV>(p, smem, s_tab, s_jlo);

    const u32 wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const u32 lane = threadIdx.x & 63;
    double dot = 0.0;   // Re<x, y> over this lane's states (used with dot_part)

    QK_FOR_TILES_W(p, wave, t, WV) {
        const XTile T = p.xtiles[t];
        const u32 hi = __builtin_amdgcn_readfirstlane(T.hi);
        if (lane >= T.nvalid) continue;
        xs_tile<SPL, NB, SYM>(p, T, hi, XRecs{p.hnb, p.hcnt, p.mnb, p.vrow}, x, y, s_tab, s_jlo, lane, dot);
    }
    if (p.dot_part) {
        // this workgroup's Re<x, y>, reduced in a fixed order
        __shared__ double s_dot[WV];
        for (int off = 32; off > 0; off >>= 1) dot += __shfl_xor(dot, off);
        if (lane == 0) s_dot[wave] = dot;
        __syncthreads();
        if (threadIdx.x == 0) {
            double tsum = 0.0;
            for (int w = 0; w < WV; ++w) tsum += s_dot[w];
            p.dot_part[blockIdx.x] = tsum;
        }
    }
}

// Multi-sector single-vector kernel: every (sector, vector) pair is a unit
// whose tiles all run on one XCD, the units spread over the 8 XCDs by size
// (the tile list holds each XCD's units in its chunk), so the vector an XCD
// works on stays in its L2 while the unit runs (the N=20 chain's largest
// sector: 2.9 MB per vector against 4 MB of L2).  Per tile the same work, in
// the same order, as k_apply_xchg_s: y is bit-identical to per-vector
// launches.  With mp.part each tile also writes Re<x, y> over its states,
// summed in a fixed order, so the recurrence's alpha needs no pass of its
// own.  WV waves per workgroup: 8 leaves a wave up to 256 registers, for
// variants that keep 16 gathers in flight.
template <int SPL, int NB, bool SYM, int WV>
__global__ void __launch_bounds__(64 * WV)
k_apply_xchg_ms(const XParams p, const MsParams mp, const uint4* __restrict__ tl,
                const cplx* __restrict__ x, cplx* __restrict__ y) {
    extern __shared__ __align__(16) unsigned char smem[];
    cplx* s_tab;
    u16* s_jlo;
    xstage_lds<WV>(p, smem, s_tab, s_jlo);

    const u32 wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const u32 lane = threadIdx.x & 63;
    // The tile list is one chunk per XCD: this wave takes every W-th entry
    // of its XCD's chunk, and loads the next entry (a scalar load from a
    // read-only argument) while the current tile runs.
    const u32 xcd = blockIdx.x % kXcd;
    const u32 W = (gridDim.x / kXcd) * WV;
    const u32 t_end = min((xcd + 1) * mp.tiles_per_xcd, mp.ntiles);
    u32 t = xcd * mp.tiles_per_xcd + (blockIdx.x / kXcd) * WV + wave;
    if (t >= t_end) return;
    uint4 e = tl[t];
    for (;;) {
        const u32 tn = t + W;
        const bool more = tn < t_end;
        const uint4 en = tl[more ? tn : t];
        const u32 unit = __builtin_amdgcn_readfirstlane(e.y);
        const u32 tile = __builtin_amdgcn_readfirstlane(e.x);
        const XTile T = mp.xt[tile];                  // T, U and S loads side by side
        const MsUnit U = mp.units[unit];
        const MsSec S = mp.secs[__builtin_amdgcn_readfirstlane(e.z)];
        const u32 hi = __builtin_amdgcn_readfirstlane(T.hi);
        double dot = 0.0;
        if (lane < T.nvalid)
            xs_tile<SPL, NB, SYM>(p, T, hi, S.rec, x + U.xoff, y + U.xoff, s_tab, s_jlo, lane, dot);
        if (mp.part && T.nvalid) {                    // padding: an empty tile
            // the tile's Re<x, y>: every lane, fixed butterfly order
            for (int off = 32; off > 0; off >>= 1) dot += __shfl_xor(dot, off);
            if (lane == 0) mp.part[U.poff + (tile - S.t0)] = dot;
        }
        if (!more) break;
        t = tn;
        e = en;
    }
}

// Exact diagonal: one group per diagonal term, summed in OpSum order with
// plain additions, as the reference accumulates h_diagonal (cpp:87-89).
template <typename S>
__global__ void __launch_bounds__(kBlock)
k_diagonal(const ApplyParams p, cplx* __restrict__ d) {
#pragma clang fp contract(off)
    const u32 wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const u32 lane = threadIdx.x & 63;
    QK_FOR_TILES(p, wave, t) {
        const uint2 td = p.tiles[t];
        const u32 hi = __builtin_amdgcn_readfirstlane(td.x);
        const u32 r0 = __builtin_amdgcn_readfirstlane(td.y);
        const int cls = p.hi_cls[hi];
        const u32 r = r0 + lane;
        if (r >= p.cls_cnt[cls]) continue;
        const u32 lo = p.lo_list[p.cls_off[cls] + r];
        const S s = (S(hi) << p.L) | S(lo);
        cplx acc{0.0, 0.0};
        for (u32 g = 0; g < p.ngroups; ++g) {
            const DevGroup G = p.groups[g];
            cplx v = p.tables[G.tab + config_of(s, G)];
            if (G.jw && (popc_s(s & S(G.jw)) & 1)) { v.re = -v.re; v.im = -v.im; }
            acc.re = acc.re + v.re;
            acc.im = acc.im + v.im;
        }
        d[p.jhi[hi] + r] = acc;
    }
}

// phi_k = N^-1/2 sum_j e^{i q_k j} Sz_j psi, q_k = 2 pi k / N.
template <typename S>
__global__ void __launch_bounds__(kBlock)
k_sz_q(const ApplyParams p, const cplx* __restrict__ phase, int nsites, int nq,
       const cplx* __restrict__ psi, cplx* __restrict__ phi, u64 ld, double scale) {
    const u32 wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const u32 lane = threadIdx.x & 63;
    QK_FOR_TILES(p, wave, t) {
        const uint2 td = p.tiles[t];
        const u32 hi = __builtin_amdgcn_readfirstlane(td.x);
        const u32 r0 = __builtin_amdgcn_readfirstlane(td.y);
        const int cls = p.hi_cls[hi];
        const u32 r = r0 + lane;
        if (r >= p.cls_cnt[cls]) continue;
        const u32 lo = p.lo_list[p.cls_off[cls] + r];
        const S s = (S(hi) << p.L) | S(lo);
        const u64 a = p.jhi[hi] + r;
        const cplx ps = psi[a];
        for (int k = 0; k < nq; ++k) {
            cplx f{0.0, 0.0};
            for (int j = 0; j < nsites; ++j) {
                const double sz = ((s >> j) & 1) ? 0.5 : -0.5;
                const cplx e = phase[k * nsites + j];
                f.re += sz * e.re;
                f.im += sz * e.im;
            }
            f.re *= scale;
            f.im *= scale;
            cplx o{0.0, 0.0};
            cfma(o, f, ps);
            phi[u64(k) * ld + a] = o;
        }
    }
}

template <typename S, int B>
void launch_apply_t(const Hamiltonian& h, const cplx* x, cplx* y, u64 ld, hipStream_t st) {
    // the exchange kernels work on (row, rank) pairs and never form a state
    // word except as 64 bits, so they serve both state widths
    if (h.xchg) {
        if constexpr (B == 1) {
            const cplx* xa = x;
            cplx* ya = y;
            void* args[] = {const_cast<XParams*>(&h.xparams_s), &xa, &ya};
            QK_HIP(hipLaunchKernel(h.xs_fn, dim3(h.xgrid_s), dim3(64 * h.xs_wv), args,
                                   size_t(h.xlds_bytes), st));
            return;
        }
        hipLaunchKernelGGL((k_apply_xchg<B>), dim3(h.xgrid), dim3(64 * kXWaves), h.xlds_bytes, st,
                           h.xparams, x, y, ld);
        return;
    }
    if (h.rows && (B == 1 ? h.rows_1 : h.rows_b)) {
        hipLaunchKernelGGL((k_apply<S, B, false, true>), dim3(h.grid), dim3(kGBlock), h.rlds_bytes, st,
                           h.rest, ColTables{}, h.rparams, x, y, ld);
        return;
    }
    hipLaunchKernelGGL((k_apply<S, B, false, false>), dim3(h.grid), dim3(kGBlock), h.lds_bytes, st,
                       h.params, ColTables{}, RowParams{}, x, y, ld);
}

template <typename S>
void launch_apply_s(const Hamiltonian& h, const cplx* x, cplx* y, int nvec, u64 ld,
                    hipStream_t st) {
    // the exchange path may run blocks as groups of fewer vectors (1: one
    // single-vector launch per vector), as chosen at set-up
    const int gmax = h.xchg ? h.vec_group : h.gen_group;
    int v = 0;
    while (v < nvec) {
        const int left = std::min(nvec - v, gmax);
        const cplx* xv = x + u64(v) * ld;
        cplx* yv = y + u64(v) * ld;
        if (left >= 8) { launch_apply_t<S, 8>(h, xv, yv, ld, st); v += 8; }
        else if (left >= 4) { launch_apply_t<S, 4>(h, xv, yv, ld, st); v += 4; }
        else if (left >= 2) { launch_apply_t<S, 2>(h, xv, yv, ld, st); v += 2; }
        else { launch_apply_t<S, 1>(h, xv, yv, ld, st); v += 1; }
    }
    QK_HIP(hipGetLastError());
}

std::vector<DevGroup> pack_groups(const std::vector<Group>& gs, std::vector<cplx>& tables) {
    std::vector<DevGroup> out;
    for (const auto& g : gs) {
        DevGroup d{};
        d.flip = g.flip;
        d.jw = g.jw;
        d.tab = static_cast<u32>(tables.size());
        d.nsup = static_cast<u32>(g.pos.size());
        d.kind = static_cast<u32>(g.kind);
        d.hi_only = g.hi_only ? 1u : 0u;
        for (size_t k = 0; k < g.pos.size(); ++k) d.posw |= u64(g.pos[k] & 63) << (6 * k);
        tables.insert(tables.end(), g.table.begin(), g.table.end());
        out.push_back(d);
    }
    return out;
}

}  // namespace

// Resident persistent grid for a kernel using `lds` bytes of LDS per
// workgroup: as many workgroups as fit on the chip at once (LDS and the
// 8-waves-per-SIMD register budget), a multiple of the XCD count.  The tile
// chunk is a few tiles per resident wave of one XCD; when that leaves fewer
// than 16 chunks per XCD, it is resized so that the chunks split evenly
// over the 8 XCDs (t-J 4x4: 10,578 tiles in chunks of 2,048 left two XCDs
// idle; a 184,756-state sector was one chunk, one XCD).
static int grid_for(int lds, size_t ntiles, u32* chunk, int waves = kWaves, int resident = 0) {
    const int cus = device_cu_count();
    const int by_lds = (160 * 1024) / std::max(lds, 1);
    const int by_waves = 32 / waves;   // 8 waves per SIMD, 4 SIMDs per CU
    int per_cu = std::max(1, std::min({8, by_lds, by_waves}));
    // a kernel whose registers let fewer workgroups reside per CU gets that
    // many when the caller passes it (the single-vector exchange kernel:
    // N=28's 68-VGPR variant holds one 16-wave workgroup per CU, and two per
    // CU ran the grid in two phases, 3 % slower).  The row-program and
    // group-loop kernels, whose tiles differ in cost, keep the second phase:
    // its workgroups start as others finish (t-J 4x4, 8 vectors: 0.209 ms
    // with it, 0.257 ms without).
    if (resident > 0) per_cu = std::min(per_cu, resident);
    // fewer resident workgroups keep fewer tiles in flight per XCD (a smaller
    // L2 working set) at the cost of latency hiding; the autotuner tries it
    if (const char* e = std::getenv("QKRYLOV_WG_PER_CU")) per_cu = std::max(1, std::min(per_cu, std::atoi(e)));
    const int want = int((ntiles + waves - 1) / waves);
    int blocks = std::min(want, cus * per_cu);
    blocks = std::max(kXcd, (blocks + kXcd - 1) / kXcd * kXcd);
    int mult = 2;
    if (const char* m = std::getenv("QKRYLOV_CHUNK_MULT")) mult = std::max(1, std::atoi(m));
    const u64 c0 = u64(std::max(64, mult * (blocks / kXcd) * waves));
    const u64 per_xcd = (ntiles + kXcd - 1) / kXcd;             // tiles per XCD
    const u64 k = std::max<u64>(1, (per_xcd + c0 - 1) / c0);    // chunks per XCD
    *chunk = u32(k >= 16 ? c0 : std::max<u64>(1, (per_xcd + k - 1) / k));
    return blocks;
}

// Workgroups of `waves` waves that one CU keeps resident for kernel fn:
// 512 VGPRs per SIMD lane shared by its waves (at most 8 waves per SIMD, 4
// SIMDs per CU) and 160 KB of LDS; 0 when the kernel's attributes are not
// available.
static int resident_wgs(const void* fn, int lds, int waves) {
    hipFuncAttributes a{};
    if (hipFuncGetAttributes(&a, fn) != hipSuccess || a.numRegs <= 0) return 0;
    const int regs = (a.numRegs + 7) / 8 * 8;   // allocation granule
    const int per_simd = std::min(8, 512 / regs);
    const int by_regs = per_simd * 4 / waves;
    const int by_lds = (160 * 1024) / std::max(lds, 1);
    return std::max(1, std::min(by_regs, by_lds));
}

int device_cu_count() {
    static int cus = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 256;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 256;
        return n > 0 ? n : 256;
    }();
    return cus;
}

Hamiltonian::Hamiltonian(std::shared_ptr<const Basis> b, SiteKind sk, const OpSum& o)
    : basis(std::move(b)), site(sk), ops(o) {
    prog = compile_program(*basis, site, ops);
    QK_HIP(hipGetDevice(&device));
    QK_HIP(hipStreamCreate(&stream));
    const Basis& B = *basis;
    state32 = B.nbits() <= 32;

    // Tiles: 64 consecutive ranks of one row, rows ordered class-major (all
    // rows whose low halves share one class, i.e. whose high halves share
    // one popcount, are adjacent: high-half flips stay inside that set).
    std::vector<u32> rows(B.valid_hi().begin(), B.valid_hi().end());
    const char* order = std::getenv("QKRYLOV_TILE_ORDER");
    if (!(order && std::string(order) == "hi"))
        std::stable_sort(rows.begin(), rows.end(), [&](u32 a, u32 b) {
            return B.hi_cls()[a] < B.hi_cls()[b];
        });
    std::vector<uint2> tiles;
    for (u32 hi : rows) {
        const u32 cnt = B.cls_cnt()[B.hi_cls()[hi]];
        for (u32 r0 = 0; r0 < cnt; r0 += 64) tiles.push_back(make_uint2(hi, r0));
    }
    std::vector<cplx> tables, dtables;
    std::vector<DevGroup> groups = pack_groups(prog.groups, tables);
    std::vector<DevGroup> dgroups = pack_groups(prog.diag_terms, dtables);

    upload(d_jlo, B.jlo().data(), B.jlo().size(), stream);
    upload(d_jhi, B.jhi().data(), B.jhi().size(), stream);
    upload(d_hicls, B.hi_cls().data(), B.hi_cls().size(), stream);
    upload(d_clsoff, B.cls_off().data(), B.cls_off().size(), stream);
    upload(d_clscnt, B.cls_cnt().data(), B.cls_cnt().size(), stream);
    upload(d_lolist, B.lo_list().data(), B.lo_list().size(), stream);
    upload(d_tiles, tiles.data(), tiles.size(), stream);
    tiles_host = tiles;
    upload(d_groups, groups.data(), groups.size(), stream);
    upload(d_tables, tables.data(), tables.size(), stream);
    upload(d_dgroups, dgroups.data(), dgroups.size(), stream);
    upload(d_dtables, dtables.data(), dtables.size(), stream);
    QK_HIP(hipStreamSynchronize(stream));

    ApplyParams& p = params;
    p.jlo = d_jlo.get<u16>();
    p.jhi = d_jhi.get<u64>();
    p.hi_cls = d_hicls.get<int>();
    p.cls_off = d_clsoff.get<u32>();
    p.cls_cnt = d_clscnt.get<u32>();
    p.lo_list = d_lolist.get<u32>();
    p.tiles = d_tiles.get<uint2>();
    p.groups = d_groups.get<DevGroup>();
    p.tables = d_tables.get<cplx>();
    p.m1lo = B.m1() & B.lomask();
    p.m2lo = B.m2() & B.lomask();
    p.ntiles = static_cast<u32>(tiles.size());
    p.ngroups = static_cast<u32>(groups.size());
    p.n_diag = prog.n_diag;
    p.n_lo = prog.n_lo;
    p.n_hi = prog.n_hi;
    p.n_mix = prog.n_mix;
    p.ntab = static_cast<u32>(tables.size());
    p.nlo = static_cast<u32>(B.jlo().size());
    p.L = B.L();
    p.ncls2 = B.ncls2();

    lds_bytes = int(p.ntab * sizeof(cplx) + ((p.nlo * sizeof(u16) + 15) / 16) * 16);
    grid = grid_for(lds_bytes, tiles.size(), &p.tiles_per_xcd, kGWaves);
    if (lds_bytes > 160 * 1024) fail("group tables exceed LDS");
    if (lds_bytes > 64 * 1024) {
        auto set = [&](const void* f) {
            QK_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes));
        };
        if (state32) {
            set((const void*)k_apply<u32, 1, false, false>); set((const void*)k_apply<u32, 2, false, false>);
            set((const void*)k_apply<u32, 4, false, false>); set((const void*)k_apply<u32, 8, false, false>);
            set((const void*)k_apply<u32, 1, true, false>); set((const void*)k_apply<u32, 2, true, false>);
            set((const void*)k_apply<u32, 4, true, false>); set((const void*)k_apply<u32, 8, true, false>);
        } else {
            set((const void*)k_apply<u64, 1, false, false>); set((const void*)k_apply<u64, 2, false, false>);
            set((const void*)k_apply<u64, 4, false, false>); set((const void*)k_apply<u64, 8, false, false>);
            set((const void*)k_apply<u64, 1, true, false>); set((const void*)k_apply<u64, 2, true, false>);
            set((const void*)k_apply<u64, 4, true, false>); set((const void*)k_apply<u64, 8, true, false>);
        }
    }

    dparams = p;
    dparams.groups = d_dgroups.get<DevGroup>();
    dparams.tables = d_dtables.get<cplx>();
    dparams.ngroups = static_cast<u32>(dgroups.size());
    dparams.ntab = static_cast<u32>(dtables.size());

    const char* mode = std::getenv("QKRYLOV_APPLY");
    force_generic = mode && std::string(mode) == "generic";
    const bool force_rows = mode && std::string(mode) == "rows";
    if (!force_generic && !force_rows) setup_exchange();
    if (!force_generic && !xchg) setup_rows();
}

// Row-program path for the groups that lie inside one half (RowParams).
// Declined (generic group loop for everything) when no group qualifies, when
// the amplitudes take more than 65,535 distinct values or when the tables
// would be unreasonably large.
void Hamiltonian::setup_rows() {
    const Basis& B = *basis;
    const int L = B.L();
    const u64 lomask = B.lomask();
    const u32 m1 = u32(B.m1() & lomask), m2 = u32(B.m2() & lomask);
    const int ncls2 = B.ncls2();
    auto inside_lo = [&](const Group& g) {
        for (int q : g.pos)
            if (q >= L) return false;
        return (g.jw & ~lomask) == 0;
    };
    auto inside_hi = [&](const Group& g) {
        for (int q : g.pos)
            if (q < L) return false;
        return (g.jw & lomask) == 0;
    };
    // the kernels' amplitude: table entry of the support configuration,
    // negated for odd Jordan-Wigner parity
    auto amp_of = [](const Group& g, u64 st) {
        u32 cf = 0;
        for (size_t k = 0; k < g.pos.size(); ++k) cf |= u32((st >> g.pos[k]) & 1) << k;
        cplx a = g.table[cf];
        if (g.jw && (__builtin_popcountll(st & g.jw) & 1)) a = cplx{-a.re, -a.im};
        return a;
    };
    auto nonzero = [](const cplx& a) { return a.re != 0.0 || a.im != 0.0; };

    std::vector<const Group*> dlo_g, dhi_g, lo_g, hi_g;
    std::vector<Group> rest_diag, rest_lo, rest_hi, rest_mix;
    for (const Group& g : prog.groups) {
        if (g.kind == kDiag && inside_lo(g)) dlo_g.push_back(&g);
        else if (g.kind == kDiag && inside_hi(g)) dhi_g.push_back(&g);
        else if (g.kind == kLo && inside_lo(g)) lo_g.push_back(&g);
        else if (g.kind == kHi && inside_hi(g)) hi_g.push_back(&g);
        else if (g.kind == kDiag) rest_diag.push_back(g);
        else if (g.kind == kLo) rest_lo.push_back(g);
        else if (g.kind == kHi) rest_hi.push_back(g);
        else rest_mix.push_back(g);
    }
    if (dlo_g.empty() && dhi_g.empty() && lo_g.empty() && hi_g.empty()) return;

    const std::vector<u32>& cls_off = B.cls_off();
    const std::vector<u32>& cls_cnt = B.cls_cnt();
    const std::vector<u32>& lo_list = B.lo_list();
    const std::vector<u16>& jlo = B.jlo();
    const std::vector<u64>& jh = B.jhi();
    const std::vector<int>& hcls = B.hi_cls();
    const size_t ncls = cls_cnt.size(), nrow = jh.size();

    // diagonal sums
    std::vector<cplx> dlo(std::max<size_t>(lo_list.size(), 1), cplx{0.0, 0.0});
    // per-row values indexed by the high half while that stays small, else
    // by the row's number among the rows that hold states (vrow)
    const bool dense = nrow * (sizeof(cplx) + sizeof(u32)) <= (size_t(64) << 20);
    std::vector<u32> vrow;
    size_t nv = nrow;
    if (!dense) {
        vrow.assign(nrow, ~0u);
        nv = 0;
        for (size_t h = 0; h < nrow; ++h)
            if (jh[h] != ~u64(0)) vrow[h] = u32(nv++);
    }
    auto slot = [&](size_t h) { return dense ? h : size_t(vrow[h]); };
    std::vector<cplx> dhi(std::max<size_t>(nv, 1), cplx{0.0, 0.0});
    for (size_t i = 0; i < lo_list.size(); ++i)
        for (const Group* g : dlo_g) {
            const cplx a = amp_of(*g, lo_list[i]);
            dlo[i] = cplx{dlo[i].re + a.re, dlo[i].im + a.im};
        }
    for (size_t h = 0; h < nrow; ++h) {
        if (jh[h] == ~u64(0)) continue;
        cplx& dv = dhi[slot(h)];
        for (const Group* g : dhi_g) {
            const cplx a = amp_of(*g, u64(h) << L);
            dv = cplx{dv.re + a.re, dv.im + a.im};
        }
    }

    // amplitude ids
    std::map<std::pair<u64, u64>, u32> ids;
    std::vector<cplx> amps;
    bool too_many = false;
    auto id_of = [&](const cplx& a) -> u32 {
        u64 kr, ki;
        std::memcpy(&kr, &a.re, 8);
        std::memcpy(&ki, &a.im, 8);
        auto it = ids.find({kr, ki});
        if (it != ids.end()) return it->second;
        if (amps.size() >= 0xFFFF) {
            too_many = true;
            return 0;
        }
        const u32 id = u32(amps.size());
        ids[{kr, ki}] = id;
        amps.push_back(a);
        return id;
    };

    // low-half groups: per class, padded column-major entries per rank
    const size_t kMaxEntries = size_t(1) << 26;
    std::vector<u32> ell, ell_off(std::max<size_t>(ncls, 1), 0u), ell_w(std::max<size_t>(ncls, 1), 0u);
    for (size_t c = 0; c < ncls && !lo_g.empty(); ++c) {
        const u32 cnt = cls_cnt[c], off = cls_off[c];
        std::vector<std::vector<u32>> per(cnt);
        u32 w = 0;
        for (u32 r = 0; r < cnt; ++r) {
            const u32 lo = lo_list[off + r];
            for (const Group* g : lo_g) {
                const cplx a = amp_of(*g, lo);
                if (!nonzero(a)) continue;
                const u32 tlo = lo ^ u32(g->flip);
                const u16 rl = jlo[tlo];
                if (rl == 0xFFFF) continue;
                if (u32(__builtin_popcount(tlo & m1) * ncls2 + __builtin_popcount(tlo & m2)) != c) continue;
                per[r].push_back(u32(rl) | (id_of(a) << 16));
            }
            w = std::max<u32>(w, u32(per[r].size()));
        }
        if (too_many) return;
        ell_off[c] = u32(ell.size());
        ell_w[c] = w;
        if (ell.size() + size_t(w) * cnt > kMaxEntries) return;
        ell.resize(ell.size() + size_t(w) * cnt, ~0u);
        for (u32 r = 0; r < cnt; ++r)
            for (size_t k = 0; k < per[r].size(); ++k) ell[ell_off[c] + k * cnt + r] = per[r][k];
    }

    // high-half groups: per row, the partner row offsets and amplitudes
    std::vector<u32> rptr(nv + 1, 0u);
    std::vector<RowEntry> rent;
    for (size_t h = 0; h < nrow; ++h) {
        if (dense) rptr[h] = u32(rent.size());
        if (jh[h] == ~u64(0)) continue;
        if (!dense) rptr[vrow[h]] = u32(rent.size());
        for (const Group* g : hi_g) {
            const size_t h2 = h ^ size_t(g->flip >> L);
            if (jh[h2] == ~u64(0) || hcls[h2] != hcls[h]) continue;
            const cplx a = amp_of(*g, u64(h) << L);
            if (!nonzero(a)) continue;
            rent.push_back(RowEntry{jh[h2], a});
        }
        if (rent.size() > kMaxEntries / 4) return;
    }
    rptr[nv] = u32(rent.size());

    // the remaining groups keep the group loop, except those across the
    // split, which become per-row entries over low-half sub-tables
    std::vector<Group> restg;
    for (auto* v : {&rest_diag, &rest_lo, &rest_hi}) restg.insert(restg.end(), v->begin(), v->end());
    std::vector<cplx> rtab;
    std::vector<DevGroup> rdev = pack_groups(restg, rtab);

    std::vector<u32> mptr;
    std::vector<RowMix> ment;
    std::vector<DevGroup> mdesc;
    if (!rest_mix.empty()) {
        struct Split {
            std::vector<int> lo_k, hi_k;   // config bits on each half
            std::vector<i64> sub;          // per high-half configuration: rtab offset or -1
            u64 fhi, jhi_bits;
        };
        std::vector<Split> sp(rest_mix.size());
        for (size_t gi = 0; gi < rest_mix.size(); ++gi) {
            const Group& g = rest_mix[gi];
            Split& S = sp[gi];
            DevGroup D{};
            for (size_t k = 0; k < g.pos.size(); ++k) {
                if (g.pos[k] < L) {
                    D.posw |= u64(g.pos[k] & 63) << (6 * S.lo_k.size());
                    S.lo_k.push_back(int(k));
                } else {
                    S.hi_k.push_back(int(k));
                }
            }
            D.flip = g.flip & lomask;
            D.jw = g.jw & lomask;
            D.nsup = u32(S.lo_k.size());
            D.kind = u32(kMix);
            mdesc.push_back(D);
            S.fhi = g.flip >> L;
            S.jhi_bits = g.jw >> L;
            const u32 nh = 1u << S.hi_k.size(), nl = 1u << S.lo_k.size();
            S.sub.assign(nh, -1);
            for (u32 ch = 0; ch < nh; ++ch) {
                std::vector<cplx> t(nl);
                bool any = false;
                for (u32 cl = 0; cl < nl; ++cl) {
                    u32 cf = 0;
                    for (size_t a = 0; a < S.lo_k.size(); ++a) cf |= ((cl >> a) & 1u) << S.lo_k[a];
                    for (size_t a = 0; a < S.hi_k.size(); ++a) cf |= ((ch >> a) & 1u) << S.hi_k[a];
                    t[cl] = g.table[cf];
                    any = any || nonzero(t[cl]);
                }
                if (!any) continue;
                S.sub[ch] = i64(rtab.size());
                rtab.insert(rtab.end(), t.begin(), t.end());
            }
        }
        mptr.assign(nv + 1, 0u);
        for (size_t h = 0; h < nrow; ++h) {
            if (dense) mptr[h] = u32(ment.size());
            if (jh[h] == ~u64(0)) continue;
            if (!dense) mptr[vrow[h]] = u32(ment.size());
            for (size_t gi = 0; gi < rest_mix.size(); ++gi) {
                const Group& g = rest_mix[gi];
                const Split& S = sp[gi];
                u32 ch = 0;
                for (size_t a = 0; a < S.hi_k.size(); ++a) ch |= u32((h >> (g.pos[S.hi_k[a]] - L)) & 1u) << a;
                if (S.sub[ch] < 0) continue;
                const size_t h2 = h ^ size_t(S.fhi);
                if (jh[h2] == ~u64(0)) continue;
                const u32 par = u32(__builtin_popcountll(u64(h) & S.jhi_bits) & 1);
                ment.push_back(RowMix{jh[h2], u32(S.sub[ch]), (u32(gi) << 1) | par, hcls[h2], 0u});
            }
            if (ment.size() > kMaxEntries / 4) return;
        }
        mptr[nv] = u32(ment.size());
    }
    const int lds = int(rtab.size() * sizeof(cplx) + ((B.jlo().size() * sizeof(u16) + 15) / 16) * 16 +
                        std::max<size_t>(amps.size(), 1) * sizeof(cplx) + mdesc.size() * sizeof(DevGroup));
    if (lds > 160 * 1024) return;

    upload(d_rgroups, rdev.data(), rdev.size(), stream);
    upload(d_rtables, rtab.data(), rtab.size(), stream);
    upload(d_rdlo, dlo.data(), dlo.size(), stream);
    upload(d_rdhi, dhi.data(), dhi.size(), stream);
    if (ell.empty()) ell.push_back(~0u);
    upload(d_ell, ell.data(), ell.size(), stream);
    upload(d_elloff, ell_off.data(), ell_off.size(), stream);
    upload(d_ellw, ell_w.data(), ell_w.size(), stream);
    if (amps.empty()) amps.push_back(cplx{0.0, 0.0});
    upload(d_amps, amps.data(), amps.size(), stream);
    upload(d_rptr, rptr.data(), rptr.size(), stream);
    if (!dense) upload(d_rvrow, vrow.data(), vrow.size(), stream);
    if (!ment.empty()) {
        upload(d_mptr, mptr.data(), mptr.size(), stream);
        upload(d_ment, ment.data(), ment.size(), stream);
        upload(d_mdesc, mdesc.data(), mdesc.size(), stream);
    }
    if (rent.empty()) rent.push_back(RowEntry{0, cplx{0.0, 0.0}});
    upload(d_rent, rent.data(), rent.size(), stream);
    QK_HIP(hipStreamSynchronize(stream));

    rest = params;
    rest.groups = d_rgroups.get<DevGroup>();
    rest.tables = d_rtables.get<cplx>();
    rest.ngroups = u32(rdev.size());
    rest.n_diag = u32(rest_diag.size());
    rest.n_lo = u32(rest_lo.size());
    rest.n_hi = u32(rest_hi.size());
    rest.n_mix = 0;
    rest.ntab = u32(rtab.size());
    RowParams& q = rparams;
    q.dlo = d_rdlo.get<cplx>();
    q.dhi = d_rdhi.get<cplx>();
    q.ell = d_ell.get<u32>();
    q.ell_off = d_elloff.get<u32>();
    q.ell_w = d_ellw.get<u32>();
    q.amps = d_amps.get<cplx>();
    q.rptr = d_rptr.get<u32>();
    q.vrow = dense ? nullptr : d_rvrow.get<u32>();
    q.mptr = ment.empty() ? nullptr : d_mptr.get<u32>();
    q.ment = ment.empty() ? nullptr : d_ment.get<RowMix>();
    q.mdesc = ment.empty() ? nullptr : d_mdesc.get<DevGroup>();
    q.nmdesc = ment.empty() ? 0u : u32(mdesc.size());
    q.rent = d_rent.get<RowEntry>();
    q.namp = u32(amps.size());
    q.ndiag = u32(dlo_g.size() + dhi_g.size());
    rlds_bytes = lds;
    if (rlds_bytes > 64 * 1024) {
        const void* f32[] = {(const void*)k_apply<u32, 1, false, true>, (const void*)k_apply<u32, 2, false, true>,
                             (const void*)k_apply<u32, 4, false, true>, (const void*)k_apply<u32, 8, false, true>};
        const void* f64[] = {(const void*)k_apply<u64, 1, false, true>, (const void*)k_apply<u64, 2, false, true>,
                             (const void*)k_apply<u64, 4, false, true>, (const void*)k_apply<u64, 8, false, true>};
        for (const void* f : (state32 ? f32 : f64))
            QK_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, rlds_bytes));
    }
    rows = true;
}

// Select the exchange fast path when every off-diagonal group is a real
// two-bit exchange and the sector counts all bits or none.
void Hamiltonian::setup_exchange() {
    const Basis& B = *basis;
    const u64 all = B.nbits() >= 64 ? ~u64(0) : ((u64(1) << B.nbits()) - 1);
    if (B.kind() != SiteKind::SpinHalf && B.kind() != SiteKind::Fermion) return;
    if (B.m2() != 0 || (B.m1() != 0 && B.m1() != all)) return;
    const int L = B.L();
    std::vector<XBond> lo_b, hi_b, mix_b;
    std::vector<const Group*> dlo_g, dhi_g, dmix_g;
    for (const Group& g : prog.groups) {
        if (g.jw != 0) return;
        for (const cplx& v : g.table)
            if (v.im != 0.0) return;
        if (g.kind == kDiag) continue;   // rebuilt per term below
        if (popc(g.flip) != 2 || g.pos.size() != 2) return;
        const int bi = g.pos[0], bj = g.pos[1];
        if (g.flip != ((u64(1) << bi) | (u64(1) << bj))) return;
        if (g.table[0].re != 0.0 || g.table[3].re != 0.0) return;
        XBond x{};
        x.bi = u32(bi);
        x.bj = u32(bj);
        x.a10 = g.table[1].re;   // config bit0 = bit(bi) = 1, bit1 = bit(bj) = 0
        x.a01 = g.table[2].re;
        x.mlo = u32(g.flip & B.lomask());
        x.mhi = u32(g.flip >> L);
        if (g.kind == kLo) lo_b.push_back(x);
        else if (g.kind == kHi) hi_b.push_back(x);
        else mix_b.push_back(x);
    }
    // Diagonal, term by term: terms on the low half go into a table aligned
    // with lo_list, terms on the high half into a per-row value; only terms
    // that straddle the split are evaluated per lane (merged by support).
    std::vector<Group> mixg;
    for (const Group& g : prog.diag_terms) {
        if (g.jw != 0) return;
        for (const cplx& v : g.table)
            if (v.im != 0.0) return;
        u64 sup = 0;
        for (int p : g.pos) sup |= u64(1) << p;
        if ((sup >> L) == 0) dlo_g.push_back(&g);
        else if ((sup & B.lomask()) == 0) dhi_g.push_back(&g);
        else {
            bool merged = false;
            for (Group& m : mixg)
                if (m.pos == g.pos) {
                    for (size_t c = 0; c < m.table.size(); ++c) m.table[c] = cadd(m.table[c], g.table[c]);
                    merged = true;
                    break;
                }
            if (!merged) mixg.push_back(g);
        }
    }
    auto eval = [](const Group& g, u64 s) {
        u32 c = 0;
        for (size_t k = 0; k < g.pos.size(); ++k) c |= u32((s >> g.pos[k]) & 1) << k;
        return g.table[c].re;
    };
    std::vector<double> dlo(B.lo_list().size(), 0.0), dhi(B.jhi().size(), 0.0);
    for (size_t i = 0; i < dlo.size(); ++i)
        for (const Group* g : dlo_g) dlo[i] += eval(*g, B.lo_list()[i]);
    for (size_t h = 0; h < dhi.size(); ++h)
        for (const Group* g : dhi_g) dhi[h] += eval(*g, u64(h) << L);
    std::vector<cplx> mixtab;
    std::vector<DevGroup> mixdev = pack_groups(mixg, mixtab);

    dhi_host = dhi;
    dlo_host = dlo;
    std::vector<XBond> bonds = lo_b;
    bonds.insert(bonds.end(), hi_b.begin(), hi_b.end());
    bonds.insert(bonds.end(), mix_b.begin(), mix_b.end());

    bonds_host = bonds;
    xsym = true;
    for (const std::vector<XBond>* v : {&lo_b, &mix_b})
        for (const XBond& b : *v) xsym = xsym && b.a01 == b.a10;
    // QKRYLOV_XSYM=0 keeps the per-lane amplitude selection (the general kernel)
    if (const char* e = std::getenv("QKRYLOV_XSYM")) xsym = xsym && std::atoi(e) != 0;
    upload(d_xbonds, bonds.data(), bonds.size(), stream);
    upload(d_dlo, dlo.data(), dlo.size(), stream);
    upload(d_dhi, dhi.data(), dhi.size(), stream);
    dmix_host = mixdev;
    dmixtab_host = mixtab;
    upload(d_dmix, mixdev.data(), mixdev.size(), stream);
    upload(d_dmixtab, mixtab.data(), mixtab.size(), stream);
    QK_HIP(hipStreamSynchronize(stream));

    // Per-row bond records of the single-vector kernel, indexed by the row
    // itself while the table for every possible high half stays small, else
    // by the row's number among the rows that hold states (vrow): a sparse
    // sector with a wide high half (few states, up to 2^26 high halves)
    // keeps records only for its own rows.
    {
        const std::vector<u64>& jh = B.jhi();
        const size_t nrow = jh.size(), nh = hi_b.size(), nm = mix_b.size();
        const size_t per_row = nh * sizeof(XRowNb) + nm * sizeof(u64) + sizeof(u32);
        const bool dense = nrow * per_row <= (size_t(256) << 20);
        std::vector<u32> vrow;
        size_t nv = nrow;
        if (!dense) {
            vrow.assign(nrow, ~0u);
            nv = 0;
            for (size_t h = 0; h < nrow; ++h)
                if (jh[h] != ~u64(0)) vrow[h] = u32(nv++);
        }
        std::vector<XRowNb> hnb(std::max<size_t>(nv * nh, 1), XRowNb{0, 0.0});
        std::vector<u32> hcnt(std::max<size_t>(nv, 1), 0u);
        std::vector<u64> mnb(std::max<size_t>(nv * nm, 1), ~u64(0));
        for (size_t h = 0; h < nrow; ++h) {
            if (jh[h] == ~u64(0)) continue;
            const size_t v = dense ? h : vrow[h];
            u32 c = 0;
            for (const XBond& X : hi_b) {
                const u32 vi = u32(h >> (X.bi - L)) & 1, vj = u32(h >> (X.bj - L)) & 1;
                if (vi == vj) continue;
                const u64 b2 = jh[h ^ X.mhi];
                if (b2 == ~u64(0)) fail("exchange path: high-half partner row missing");
                hnb[v * nh + c++] = XRowNb{b2, vi ? X.a10 : X.a01};
            }
            hcnt[v] = c;
            for (size_t m = 0; m < nm; ++m) mnb[v * nm + m] = jh[h ^ mix_b[m].mhi];
        }
        upload(d_hnb, hnb.data(), hnb.size(), stream);
        upload(d_hcnt, hcnt.data(), hcnt.size(), stream);
        upload(d_mnb, mnb.data(), mnb.size(), stream);
        if (!dense) upload(d_vrow, vrow.data(), vrow.size(), stream);
        xparams.vrow = dense ? nullptr : d_vrow.get<u32>();
    }

    XParams& x = xparams;
    x.hnb = d_hnb.get<XRowNb>();
    x.hcnt = d_hcnt.get<u32>();
    x.mnb = d_mnb.get<u64>();
    x.jlo = params.jlo;
    x.jhi = params.jhi;
    x.hi_cls = params.hi_cls;
    x.cls_off = params.cls_off;
    x.cls_cnt = params.cls_cnt;
    x.lo_list = params.lo_list;
    x.dlo = d_dlo.get<double>();
    x.dhi = d_dhi.get<double>();
    x.tiles = params.tiles;
    x.bonds = d_xbonds.get<XBond>();
    x.dmix = d_dmix.get<DevGroup>();
    x.dmix_tab = d_dmixtab.get<cplx>();
    x.n_lo = u32(lo_b.size());
    x.n_hi = u32(hi_b.size());
    x.n_mix = u32(mix_b.size());
    x.n_dmix = u32(mixdev.size());
    x.ndmix_tab = u32(mixtab.size());
    x.nlo = params.nlo;
    x.L = L;
    xlds_bytes = int(x.ndmix_tab * sizeof(cplx) + ((x.nlo * sizeof(u16) + 15) / 16) * 16);
    const std::vector<XTile> xt = xtiles_for(64);
    xgrid = grid_for(xlds_bytes, xt.size(), &x.tiles_per_xcd, kXWaves);
    x.ntiles = u32(xt.size());
    upload(d_xtiles, xt.data(), xt.size(), stream);
    x.xtiles = d_xtiles.get<XTile>();
    QK_HIP(hipStreamSynchronize(stream));
    if (xlds_bytes > 64 * 1024) {
        for (const void* f : {(const void*)k_apply_xchg<1>, (const void*)k_apply_xchg<2>,
                              (const void*)k_apply_xchg<4>, (const void*)k_apply_xchg<8>})
            QK_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, xlds_bytes));
    }
    // defaults (measured on the N=28 chain); autotune() may change them
    int s0 = 2, nb0 = 4;
    vec_group = 2;
    if (const char* e = std::getenv("QKRYLOV_SPL")) s0 = std::atoi(e);
    if (const char* e = std::getenv("QKRYLOV_BOND_BATCH")) nb0 = std::atoi(e);
    if (const char* e = std::getenv("QKRYLOV_VEC_GROUP")) vec_group = std::atoi(e);
    if (vec_group != 1 && vec_group != 2 && vec_group != 4 && vec_group != 8) vec_group = 2;
    xchg = true;
    set_single_kernel(s0, nb0);
    xchg = true;
}

// Tiles of `span` consecutive ranks of one row, rows in the order of
// tiles_host (class-major).
std::vector<XTile> Hamiltonian::xtiles_for(u32 span) const {
    const Basis& B = *basis;
    std::vector<XTile> xt;
    u32 last = ~0u;
    for (const uint2& tl : tiles_host) {
        const u32 hi = tl.x;
        if (hi == last) continue;
        last = hi;
        const int c = B.hi_cls()[hi];
        const u32 cnt = B.cls_cnt()[c];
        for (u32 r0 = 0; r0 < cnt; r0 += span) {
            XTile T{};
            T.base = B.jhi()[hi];
            T.pos0 = B.cls_off()[c] + r0;
            T.hi = hi;
            T.r0 = r0;
            T.nvalid = std::min<u32>(span, cnt - r0);
            T.dh = dhi_host[hi];
            xt.push_back(T);
        }
    }
    return xt;
}

// Single-vector exchange kernel with s states per lane and bonds in batches
// of nb (xbonds_*): tiles of 64*s ranks, grid, LDS limit.
void Hamiltonian::set_single_kernel(int s, int nb, int wv) {
    if (!xchg) return;
    if (s != 1 && s != 2 && s != 4) s = 2;
    if (nb != 1 && nb != 2 && nb != 4) nb = 1;
    if (s == 4 && nb == 4) nb = 2;   // that pair spills registers
    if (s != spl || !d_xtiles_s.get()) {
        const std::vector<XTile> xts = xtiles_for(64u * u32(s));
        xparams_s = xparams;
        xparams_s.ntiles = u32(xts.size());
        upload(d_xtiles_s, xts.data(), xts.size(), stream);
        xparams_s.xtiles = d_xtiles_s.get<XTile>();
        QK_HIP(hipStreamSynchronize(stream));
    }
    const void* table[2][3][3] = {
        {{(const void*)k_apply_xchg_s<1, 1, false>, (const void*)k_apply_xchg_s<1, 2, false>,
          (const void*)k_apply_xchg_s<1, 4, false>},
         {(const void*)k_apply_xchg_s<2, 1, false>, (const void*)k_apply_xchg_s<2, 2, false>,
          (const void*)k_apply_xchg_s<2, 4, false>},
         {(const void*)k_apply_xchg_s<4, 1, false>, (const void*)k_apply_xchg_s<4, 2, false>, nullptr}},
        {{(const void*)k_apply_xchg_s<1, 1, true>, (const void*)k_apply_xchg_s<1, 2, true>,
          (const void*)k_apply_xchg_s<1, 4, true>},
         {(const void*)k_apply_xchg_s<2, 1, true>, (const void*)k_apply_xchg_s<2, 2, true>,
          (const void*)k_apply_xchg_s<2, 4, true>},
         {(const void*)k_apply_xchg_s<4, 1, true>, (const void*)k_apply_xchg_s<4, 2, true>, nullptr}}};
    const auto slot = [](int v) { return v == 1 ? 0 : v == 2 ? 1 : 2; };
    spl = s;
    bond_batch = nb;
    xs_fn = table[xsym ? 1 : 0][slot(s)][slot(nb)];
    xs_wv = kXWaves;
    // 8-wave workgroups where built (2 states per lane, 2 or 4 bonds per
    // batch): by default below 2^18 states, where launches are short and a
    // finer grid fills the chip sooner; QKRYLOV_XS_WAVES sets 8 or 16
    if (wv == 0) {
        wv = dim() < (u64(1) << 18) ? 8 : 16;
        if (const char* e = std::getenv("QKRYLOV_XS_WAVES")) wv = std::atoi(e);
    }
    if (wv == 8 && s == 2 && nb != 1) {
        const void* t8[2][2] = {
            {(const void*)k_apply_xchg_s<2, 2, false, 8>, (const void*)k_apply_xchg_s<2, 4, false, 8>},
            {(const void*)k_apply_xchg_s<2, 2, true, 8>, (const void*)k_apply_xchg_s<2, 4, true, 8>}};
        xs_fn = t8[xsym ? 1 : 0][nb == 2 ? 0 : 1];
        xs_wv = 8;
    }
    if (xlds_bytes > 64 * 1024)
        QK_HIP(hipFuncSetAttribute(xs_fn, hipFuncAttributeMaxDynamicSharedMemorySize, xlds_bytes));
    // the grid: workgroups this variant keeps resident per CU (registers, LDS)
    xs_res = resident_wgs(xs_fn, xlds_bytes, xs_wv);
    xgrid_s = grid_for(xlds_bytes, xparams_s.ntiles, &xparams_s.tiles_per_xcd, xs_wv, xs_res);
}

// Autotuner: times every single-vector variant (states per lane x bond
// batch) and every block grouping on this Hamiltonian with the caller's
// block width, keeps the fastest, and returns the timings as JSON.
std::string Hamiltonian::autotune(int nvec) {
    std::ostringstream os;
    os << "{";
    if (dim() == 0 || (!xchg && !rows)) {
        os << "\"tuned\": 0}";
        return os.str();
    }
    nvec = std::max(1, std::min(nvec, 8));
    const u64 d = dim();
    DevBuf x(sizeof(cplx) * d * nvec), y(sizeof(cplx) * d * nvec);
    QK_HIP(hipMemsetAsync(x.get(), 0, sizeof(cplx) * d * nvec, stream));
    hipEvent_t e0, e1;
    QK_HIP(hipEventCreate(&e0));
    QK_HIP(hipEventCreate(&e1));
    auto time_ms = [&](int nv, int reps) {
        apply(x.get<cplx>(), y.get<cplx>(), nv, d, stream);
        QK_HIP(hipEventRecord(e0, stream));
        for (int r = 0; r < reps; ++r) apply(x.get<cplx>(), y.get<cplx>(), nv, d, stream);
        QK_HIP(hipEventRecord(e1, stream));
        QK_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        QK_HIP(hipEventElapsedTime(&ms, e0, e1));
        return double(ms) / reps;
    };
    const int reps = d > (u64(1) << 24) ? 3 : 10;
    if (!xchg) {
        // row-program against group-loop kernel, then vectors per launch
        double best1 = 1e300, bestb = 1e300;
        bool r1 = rows_1, rb = rows_b;
        int gg = gen_group;
        os << "\"single\": {";
        for (int k = 0; k < 2; ++k) {
            rows_1 = k == 0;
            const double ms = time_ms(1, reps);
            os << (k ? ", " : "") << (k == 0 ? "\"rows\"" : "\"loop\"") << ": " << ms;
            if (ms < best1) {
                best1 = ms;
                r1 = rows_1;
            }
        }
        rows_1 = r1;
        os << "}, \"row_single\": " << (r1 ? 1 : 0);
        if (nvec > 1) {
            os << ", \"groups\": {";
            bool first = true;
            for (int k = 0; k < 2; ++k)
                for (int g : {1, 2, 4, 8}) {
                    if (g > nvec && g != 1) continue;
                    if (g == 1 && k == 1) continue;   // single launches follow row_single
                    rows_b = k == 0;
                    gen_group = g;
                    const double ms = time_ms(nvec, reps);
                    os << (first ? "" : ", ") << "\"" << (k == 0 ? "rows" : "loop") << "/" << g << "\": " << ms;
                    first = false;
                    if (ms < bestb) {
                        bestb = ms;
                        rb = rows_b;
                        gg = g;
                    }
                }
            rows_b = rb;
            gen_group = gg;
            os << "}, \"row_block\": " << (rb ? 1 : 0) << ", \"gen_group\": " << gg;
        }
        os << ", \"tuned\": 1}";
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return os.str();
    }
    double best = 1e300;
    int bs = spl, bn = bond_batch, bw = xs_wv;
    os << "\"single\": {";
    bool first = true;
    for (int s : {1, 2, 4})
        for (int nb : {1, 2, 4})
            for (int wv : {16, 8}) {
                if (s == 4 && nb == 4) continue;
                if (wv == 8 && !(s == 2 && nb != 1)) continue;   // 8-wave variants built
                set_single_kernel(s, nb, wv);
                const double ms = time_ms(1, reps);
                os << (first ? "" : ", ") << "\"" << s << "x" << nb << (wv == 8 ? "w8" : "") << "\": " << ms;
                first = false;
                if (ms < best) {
                    best = ms;
                    bs = s;
                    bn = nb;
                    bw = wv;
                }
            }
    set_single_kernel(bs, bn, bw);
    os << "}, \"states_per_lane\": " << bs << ", \"bond_batch\": " << bn << ", \"waves\": " << bw;
    if (nvec > 1) {
        double bestv = 1e300;
        int bg = vec_group;
        os << ", \"groups\": {";
        first = true;
        for (int g : {1, 2, 4, 8}) {
            if (g > nvec && g != 1) continue;
            vec_group = g;
            const double ms = time_ms(nvec, reps);
            os << (first ? "" : ", ") << "\"" << g << "\": " << ms;
            first = false;
            if (ms < bestv) {
                bestv = ms;
                bg = g;
            }
        }
        vec_group = bg;
        os << "}, \"vec_group\": " << bg;
    }
    os << ", \"tuned\": 1}";
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return os.str();
}

bool MultiApply::build(const std::vector<Group>& groups, hipStream_t st) {
    ok_ = false;
    h0_ = nullptr;
    std::vector<const Group*> gs;
    for (const Group& g : groups)
        if (g.h && g.nvec > 0 && g.h->dim() > 0) gs.push_back(&g);
    if (gs.empty()) return false;
    const Hamiltonian& h0 = *gs.front()->h;
    if (!h0.xchg) return false;
    for (const Group* g : gs) {
        const Hamiltonian& h = *g->h;
        if (!h.xchg || h.basis->L() != h0.basis->L() || h.basis->nbits() != h0.basis->nbits() ||
            h.xparams.nlo != h0.xparams.nlo || h.xparams.n_dmix != h0.xparams.n_dmix ||
            h.xparams.ndmix_tab != h0.xparams.ndmix_tab || h.xlds_bytes != h0.xlds_bytes ||
            h.dlo_host != h0.dlo_host || h.dhi_host != h0.dhi_host ||
            h.bonds_host.size() != h0.bonds_host.size() || h.xparams.n_lo != h0.xparams.n_lo ||
            h.xparams.n_hi != h0.xparams.n_hi || h.xparams.n_mix != h0.xparams.n_mix ||
            std::memcmp(h.bonds_host.data(), h0.bonds_host.data(), sizeof(XBond) * h.bonds_host.size()) != 0 ||
            h.dmix_host.size() != h0.dmix_host.size() || h.dmixtab_host.size() != h0.dmixtab_host.size() ||
            std::memcmp(h.dmix_host.data(), h0.dmix_host.data(), sizeof(DevGroup) * h.dmix_host.size()) != 0 ||
            std::memcmp(h.dmixtab_host.data(), h0.dmixtab_host.data(), sizeof(cplx) * h.dmixtab_host.size()) != 0)
            return false;
    }
    h0_ = &h0;
    // QKRYLOV_MULTI_KERNEL=block: the block kernel, every tile applied to
    // several vectors of its sector; default: one vector per tile, each
    // vector's tiles on one XCD
    const char* mk = std::getenv("QKRYLOV_MULTI_KERNEL");
    if (!(mk && std::string(mk) == "block") && build_units(gs, st)) {
        units_ = true;
        ok_ = true;
        return true;
    }
    units_ = false;
    std::vector<MTile> mt;
    std::vector<const u64*> jh;
    // vblk > 0: one tile per (block of vblk vectors, tile), all tiles of the
    // first block first, so the grid gathers from one block of vectors at a
    // time and its rows stay in L2 between neighbouring tiles (N=20 FTLM,
    // 64 vectors per sector: 8 per block 0.436 s per sweep against 0.551 s
    // for all 64 per tile; 4: 0.448 s, 16: 0.457 s, 32: 0.467 s); 0: one
    // tile with all of a group's vectors
    int vblk = 8;
    if (const char* e = std::getenv("QKRYLOV_MULTI_VBLOCK")) vblk = std::atoi(e);
    int maxv = 0;
    for (const Group* g : gs) maxv = std::max(maxv, g->nvec);
    const int nvb = vblk > 0 ? (maxv + vblk - 1) / vblk : 1;
    for (size_t s = 0; s < gs.size(); ++s) jh.push_back(gs[s]->h->xparams.jhi);
    for (int b = 0; b < nvb; ++b)
        for (size_t s = 0; s < gs.size(); ++s) {
            const Group& g = *gs[s];
            const int v0 = vblk > 0 ? b * vblk : 0;
            const int nv = vblk > 0 ? std::min(vblk, g.nvec - v0) : g.nvec;
            if (nv <= 0) continue;
            for (const XTile& T : g.h->xtiles_for(64)) {
                MTile m{};
                m.base = T.base;
                m.xoff = g.off + u64(v0) * g.h->dim();
                m.ld = g.h->dim();
                m.pos0 = T.pos0;
                m.hi = T.hi;
                m.r0 = T.r0;
                m.nvalid = T.nvalid;
                m.sec = u32(s);
                m.nvec = u32(nv);
                m.dh = T.dh;
                mt.push_back(m);
            }
        }
    upload(tiles_, mt.data(), mt.size(), st);
    upload(jhis_, jh.data(), jh.size(), st);
    QK_HIP(hipStreamSynchronize(st));
    mp_.tiles = tiles_.get<MTile>();
    mp_.jhis = jhis_.get<const u64*>();
    mp_.ntiles = u32(mt.size());
    grid_ = grid_for(h0.xlds_bytes, mt.size(), &mp_.tiles_per_xcd, kXWaves);
    vb_ = 8;   // vectors per pass over a tile
    if (const char* e = std::getenv("QKRYLOV_MULTI_B")) vb_ = std::atoi(e) == 4 ? 4 : 8;
    if (h0.xlds_bytes > 64 * 1024) {
        QK_HIP(hipFuncSetAttribute((const void*)k_apply_xchg_multi<4>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, h0.xlds_bytes));
        QK_HIP(hipFuncSetAttribute((const void*)k_apply_xchg_multi<8>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, h0.xlds_bytes));
    }
    h0_ = &h0;
    ok_ = true;
    return true;
}

// Units: one (group, vector) pair each, costed by its tiles.  With fewer
// than 32 units each is cut into pieces of consecutive tiles, so that a
// single vector still spreads over every XCD.  Each piece goes to the XCD
// with the least work so far, largest first; an XCD's list holds its pieces
// one after another and is the XCD's one chunk of the tile schedule (chunk c
// runs on XCD c % 8), padded with empty entries to the longest list.
bool MultiApply::build_units(const std::vector<const Group*>& gs, hipStream_t st) {
    const Hamiltonian& h0 = *h0_;
    for (const Group* g : gs)
        if (g->h->xsym != h0.xsym) return false;
    // (states per lane, bonds per batch): up to 16 gathers per wave in flight
    // (states per lane, bonds per batch, waves per workgroup): up to 16
    // gathers per wave in flight
    struct Variant {
        int s, nb, wv;
        const void* fn[2];   // [SYM]
    };
#define QK_MS(S, NB, WV)                                                                    \
    Variant {                                                                               \
        S, NB, WV, {                                                                        \
            (const void*)k_apply_xchg_ms<S, NB, false, WV>, (const void*)k_apply_xchg_ms<S, NB, true, WV> \
        }                                                                                   \
    }
    static const Variant variants[] = {QK_MS(1, 2, 16), QK_MS(1, 4, 16), QK_MS(1, 8, 16), QK_MS(2, 1, 16),
                                       QK_MS(2, 2, 16), QK_MS(2, 4, 16), QK_MS(2, 4, 8),  QK_MS(2, 8, 8),
                                       QK_MS(4, 2, 16), QK_MS(4, 4, 8)};
#undef QK_MS
    // default (2, 4, 8): N=20 FTLM sweep 4,214 samples/s against 4,075-4,092
    // with 16-wave workgroups; 16 gathers in flight per wave ran slower
    // ((2, 8, 8) 3,250, (4, 4, 8) 3,119, (1, 8, 16) 3,388; (4, 2, 16) 4,013)
    int s = 2, nb = 4, wv = 8;
    if (const char* e = std::getenv("QKRYLOV_MULTI_SPL")) s = std::atoi(e);
    if (const char* e = std::getenv("QKRYLOV_MULTI_NB")) nb = std::atoi(e);
    if (const char* e = std::getenv("QKRYLOV_MULTI_WAVES")) wv = std::atoi(e);
    ms_fn_ = nullptr;
    for (const Variant& v : variants)
        if (v.s == s && v.nb == nb && (v.wv == wv || !ms_fn_)) {
            ms_fn_ = v.fn[h0.xsym ? 1 : 0];
            ms_wv_ = v.wv;
        }
    if (!ms_fn_) {   // not built: the default
        s = 2;
        ms_wv_ = 8;
        ms_fn_ = h0.xsym ? (const void*)k_apply_xchg_ms<2, 4, true, 8> : (const void*)k_apply_xchg_ms<2, 4, false, 8>;
    }
    if (h0.xlds_bytes > 64 * 1024)
        QK_HIP(hipFuncSetAttribute(ms_fn_, hipFuncAttributeMaxDynamicSharedMemorySize, h0.xlds_bytes));

    // per-sector tiles and records
    std::vector<XTile> xt;
    std::vector<size_t> xoff;
    std::vector<u32> ntile;
    for (const Group* g : gs) {
        const std::vector<XTile> t = g->h->xtiles_for(64u * u32(s));
        xoff.push_back(xt.size());
        ntile.push_back(u32(t.size()));
        xt.insert(xt.end(), t.begin(), t.end());
    }
    // the last tile is empty: the padding entries of short XCD lists
    const u32 empty = u32(xt.size());
    xt.push_back(XTile{});
    if (xt.size() >= (size_t(1) << 32)) return false;
    upload(ms_xt_, xt.data(), xt.size(), st);
    std::vector<MsSec> secs;
    for (size_t i = 0; i < gs.size(); ++i) {
        const XParams& q = gs[i]->h->xparams;
        secs.push_back(MsSec{XRecs{q.hnb, q.hcnt, q.mnb, q.vrow}, u32(xoff[i]), 0u});
    }
    std::vector<MsUnit> units;
    utiles_.clear();
    u32 maxt = 0;
    for (size_t i = 0; i < gs.size(); ++i) maxt = std::max(maxt, ntile[i]);
    for (size_t i = 0; i < gs.size(); ++i)
        for (int v = 0; v < gs[i]->nvec; ++v) {
            units.push_back(MsUnit{gs[i]->off + u64(v) * gs[i]->h->dim(), u32(i), 0u});
            utiles_.push_back(ntile[i]);
        }
    // tile partials: unit u at u * maxt (when the offsets fit 32 bits)
    pstride_ = u64(units.size()) * maxt < (u64(1) << 32) ? maxt : 0u;
    for (size_t u = 0; u < units.size(); ++u) units[u].poff = u32(u * pstride_);

    // pieces (unit, first tile, end tile) and the XCD lists
    struct Piece {
        u32 unit, t0, t1;
    };
    const u32 npiece = units.size() >= 4 * kXcd ? 1u : u32((4 * kXcd + units.size() - 1) / units.size());
    std::vector<Piece> pieces;
    for (u32 u = 0; u < units.size(); ++u) {
        const u32 nt = ntile[units[u].sec];
        const u32 np = std::max(1u, std::min(npiece, nt));
        for (u32 k = 0; k < np; ++k) pieces.push_back(Piece{u, u32(u64(nt) * k / np), u32(u64(nt) * (k + 1) / np)});
    }
    std::stable_sort(pieces.begin(), pieces.end(),
                     [](const Piece& a, const Piece& b) { return a.t1 - a.t0 > b.t1 - b.t0; });
    std::vector<std::vector<Piece>> xu(kXcd);
    std::vector<u64> load(kXcd, 0);
    for (const Piece& q : pieces) {
        const int j = int(std::min_element(load.begin(), load.end()) - load.begin());
        xu[j].push_back(q);
        load[j] += q.t1 - q.t0;
    }
    // grid: the workgroups the kernel keeps resident, or fewer for few tiles
    // (one wave per tile at most: the plain loop's small sectors)
    const int per_cu = std::max(1, resident_wgs(ms_fn_, h0.xlds_bytes, ms_wv_));
    const int by_lds = (160 * 1024) / std::max(h0.xlds_bytes, 1);
    const int cus = device_cu_count();
    u64 real = 0;
    for (u32 u = 0; u < units.size(); ++u) real += ntile[units[u].sec];
    const u64 want = (real + ms_wv_ - 1) / ms_wv_;
    grid_ = int(std::min<u64>(want, u64(cus) * std::max(1, std::min({per_cu, by_lds, 32 / ms_wv_}))));
    grid_ = std::max(kXcd, (grid_ + kXcd - 1) / kXcd * kXcd);
    std::vector<std::vector<uint4>> lists(kXcd);
    size_t longest = 0;
    for (int j = 0; j < kXcd; ++j) {
        std::sort(xu[j].begin(), xu[j].end(), [](const Piece& a, const Piece& b) {
            return a.unit != b.unit ? a.unit < b.unit : a.t0 < b.t0;
        });
        for (const Piece& q : xu[j]) {
            const u32 b = u32(xoff[units[q.unit].sec]);
            for (u32 t = q.t0; t < q.t1; ++t) lists[j].push_back(make_uint4(b + t, q.unit, units[q.unit].sec, 0u));
        }
        longest = std::max(longest, lists[j].size());
    }
    // one chunk per XCD: the lists back to back, each padded to the longest
    const size_t chunk = std::max<size_t>(longest, 1);
    if (chunk * kXcd >= (size_t(1) << 32)) return false;
    std::vector<uint4> tl;
    tl.reserve(chunk * kXcd);
    for (int j = 0; j < kXcd; ++j)
        for (size_t i = 0; i < chunk; ++i)
            tl.push_back(i < lists[j].size() ? lists[j][i] : make_uint4(empty, 0u, 0u, 0u));
    upload(ms_tiles_, tl.data(), tl.size(), st);
    upload(ms_units_, units.data(), units.size(), st);
    upload(ms_secs_, secs.data(), secs.size(), st);
    QK_HIP(hipStreamSynchronize(st));
    msp_.tiles = ms_tiles_.get<uint4>();
    msp_.xt = ms_xt_.get<XTile>();
    msp_.units = ms_units_.get<MsUnit>();
    msp_.secs = ms_secs_.get<MsSec>();
    msp_.part = nullptr;
    msp_.ntiles = u32(tl.size());
    msp_.tiles_per_xcd = u32(chunk);
    return true;
}

void MultiApply::launch(const cplx* x, cplx* y, hipStream_t st, double* part) const {
    if (!ok_) fail("MultiApply::launch without a successful build");
    if (part && !tile_parts()) fail("MultiApply::launch: tile partial sums need the per-vector kernel");
    if (units_) {
        XParams p = h0_->xparams;
        MsParams m = msp_;
        m.part = part;
        const cplx* xa = x;
        cplx* ya = y;
        const uint4* tl = msp_.tiles;
        void* args[] = {&p, &m, &tl, &xa, &ya};
        QK_HIP(hipLaunchKernel(ms_fn_, dim3(grid_), dim3(64 * ms_wv_), args, size_t(h0_->xlds_bytes), st));
        return;
    }
    if (vb_ == 4)
        hipLaunchKernelGGL((k_apply_xchg_multi<4>), dim3(grid_), dim3(64 * kXWaves), h0_->xlds_bytes,
                           st, h0_->xparams, mp_, x, y);
    else
        hipLaunchKernelGGL((k_apply_xchg_multi<8>), dim3(grid_), dim3(64 * kXWaves), h0_->xlds_bytes,
                           st, h0_->xparams, mp_, x, y);
    QK_HIP(hipGetLastError());
}

Hamiltonian::~Hamiltonian() {
    if (stream) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamDestroy(stream);
    }
}

bool Hamiltonian::apply_dot(const cplx* x, cplx* y, double* part, hipStream_t st) const {
    if (!xchg || !xs_fn || dim() == 0) return false;
    if (!st) st = stream;
    XParams q = xparams_s;
    q.dot_part = part;
    const cplx* xa = x;
    cplx* ya = y;
    void* args[] = {&q, &xa, &ya};
    QK_HIP(hipLaunchKernel(xs_fn, dim3(xgrid_s), dim3(64 * xs_wv), args, size_t(xlds_bytes), st));
    return true;
}

void Hamiltonian::apply(const cplx* x, cplx* y, int nvec, u64 ld, hipStream_t st) const {
    if (dim() == 0 || nvec <= 0) return;
    if (!st) st = stream;
    if (state32) launch_apply_s<u32>(*this, x, y, nvec, ld, st);
    else launch_apply_s<u64>(*this, x, y, nvec, ld, st);
}

// ---------------------------------------------------------------------------
// Operator between bases
// ---------------------------------------------------------------------------
OpSum adjoint(const OpSum& ops) {
    static const std::pair<const char*, const char*> swap[] = {
        {"Sp", "Sm"}, {"C", "Cdag"}, {"CUp", "CdagUp"}, {"CDn", "CdagDn"}};
    OpSum out;
    for (const OpTerm& t : ops.terms) {
        OpTerm a;
        a.coeff = cplx{t.coeff.re, -t.coeff.im};
        for (auto it = t.factors.rbegin(); it != t.factors.rend(); ++it) {
            OpFactor f = *it;
            for (const auto& sw : swap) {
                if (f.op == sw.first) { f.op = sw.second; break; }
                if (f.op == sw.second) { f.op = sw.first; break; }
            }
            a.factors.push_back(f);
        }
        out.terms.push_back(std::move(a));
    }
    return out;
}

Operator::Operator(std::shared_ptr<const Basis> in, std::shared_ptr<const Basis> out, SiteKind site,
                   const OpSum& ops)
    : in_(std::move(in)) {
    const Basis& A = *in_;
    const Basis& B = *out;
    if (A.kind() != B.kind() || A.nsites() != B.nsites() || A.L() != B.L() || A.m1() != B.m1() ||
        A.m2() != B.m2() || A.ncls2() != B.ncls2())
        fail("operator: input and output bases must share site kind, size and sector type");
    rows_ = std::make_unique<Hamiltonian>(std::move(out), site, adjoint(ops));
    upload(jhi_in_, A.jhi().data(), A.jhi().size(), rows_->stream);
    upload(hicls_in_, A.hi_cls().data(), A.hi_cls().size(), rows_->stream);
    QK_HIP(hipStreamSynchronize(rows_->stream));
}

template <typename S>
static void launch_operator(const Hamiltonian& h, const ColTables& c, const cplx* x, cplx* y, int nvec,
                            u64 ld_in, u64 ld_out, hipStream_t st) {
    int v = 0;
    while (v < nvec) {
        const int left = nvec - v;
        const cplx* xv = x ? x + u64(v) * ld_in : nullptr;
        cplx* yv = y + u64(v) * ld_out;
        if (left >= 8) {
            hipLaunchKernelGGL((k_apply<S, 8, true, false>), dim3(h.grid), dim3(kGBlock), h.lds_bytes, st, h.params, c, RowParams{}, xv, yv, ld_in);
            v += 8;
        } else if (left >= 4) {
            hipLaunchKernelGGL((k_apply<S, 4, true, false>), dim3(h.grid), dim3(kGBlock), h.lds_bytes, st, h.params, c, RowParams{}, xv, yv, ld_in);
            v += 4;
        } else if (left >= 2) {
            hipLaunchKernelGGL((k_apply<S, 2, true, false>), dim3(h.grid), dim3(kGBlock), h.lds_bytes, st, h.params, c, RowParams{}, xv, yv, ld_in);
            v += 2;
        } else {
            hipLaunchKernelGGL((k_apply<S, 1, true, false>), dim3(h.grid), dim3(kGBlock), h.lds_bytes, st, h.params, c, RowParams{}, xv, yv, ld_in);
            v += 1;
        }
    }
    QK_HIP(hipGetLastError());
}

void Operator::apply(const cplx* x, cplx* y, int nvec, u64 ld_in, u64 ld_out, hipStream_t st) const {
    if (dim_out() == 0 || nvec <= 0) return;
    if (!st) st = rows_->stream;
    // targets outside `in` (every target when it is empty) are skipped by the
    // row tables, so x is never read then
    const ColTables c{jhi_in_.get<u64>(), hicls_in_.get<int>(), ld_out};
    if (rows_->state32) launch_operator<u32>(*rows_, c, x, y, nvec, ld_in, ld_out, st);
    else launch_operator<u64>(*rows_, c, x, y, nvec, ld_in, ld_out, st);
}

void Hamiltonian::diagonal(cplx* d, hipStream_t st) const {
    if (dim() == 0) return;
    if (!st) st = stream;
    if (state32)
        hipLaunchKernelGGL(k_diagonal<u32>, dim3(grid), dim3(kBlock), 0, st, dparams, d);
    else
        hipLaunchKernelGGL(k_diagonal<u64>, dim3(grid), dim3(kBlock), 0, st, dparams, d);
    QK_HIP(hipGetLastError());
}

void Hamiltonian::sz_q(const cplx* psi, cplx* phi, int nq, u64 ld, hipStream_t st) const {
    if (basis->kind() != SiteKind::SpinHalf) fail("sz_q: spin-1/2 bases only");
    if (dim() == 0 || nq <= 0) return;
    if (!st) st = stream;
    const int n = basis->nsites();
    std::vector<cplx> ph(size_t(nq) * n);
    for (int k = 0; k < nq; ++k)
        for (int j = 0; j < n; ++j) {
            const double ang = 2.0 * M_PI * double(k) * double(j) / double(n);
            ph[size_t(k) * n + j] = cplx{std::cos(ang), std::sin(ang)};
        }
    DevBuf dph;
    upload(dph, ph.data(), ph.size(), st);
    const double scale = 1.0 / std::sqrt(double(n));
    if (state32)
        hipLaunchKernelGGL(k_sz_q<u32>, dim3(grid), dim3(kBlock), 0, st, params,
                           dph.get<cplx>(), n, nq, psi, phi, ld, scale);
    else
        hipLaunchKernelGGL(k_sz_q<u64>, dim3(grid), dim3(kBlock), 0, st, params,
                           dph.get<cplx>(), n, nq, psi, phi, ld, scale);
    QK_HIP(hipGetLastError());
    QK_HIP(hipStreamSynchronize(st));   // dph is freed on return
}

std::string Hamiltonian::info() const {
    std::ostringstream os;
    os << "{\"dim\": " << dim() << ", \"nbits\": " << basis->nbits() << ", \"L\": " << basis->L()
       << ", \"H\": " << basis->H() << ", \"terms\": " << ops.terms.size()
       << ", \"groups\": " << prog.groups.size() << ", \"diag\": " << prog.n_diag
       << ", \"lo\": " << prog.n_lo << ", \"hi\": " << prog.n_hi << ", \"mix\": " << prog.n_mix
       << ", \"tiles\": " << params.ntiles << ", \"grid\": " << grid
       << ", \"lds\": " << lds_bytes << ", \"state32\": " << (state32 ? 1 : 0)
       << ", \"max_support\": " << prog.max_support << ", \"exchange_path\": " << (xchg ? 1 : 0)
       << ", \"xbonds\": [" << xparams.n_lo << ", " << xparams.n_hi << ", " << xparams.n_mix
       << "], \"diag_mixed\": " << xparams.n_dmix << ", \"xsym\": " << (xsym ? 1 : 0)
       << ", \"resident_single\": " << xs_res
       << ", \"states_per_lane\": " << spl
       << ", \"bond_batch\": " << bond_batch << ", \"vec_group\": " << vec_group << ", \"xgrid\": " << xgrid_s
       << ", \"row_path\": " << (rows ? 1 : 0) << ", \"row_single\": " << (rows && rows_1 ? 1 : 0)
       << ", \"row_block\": " << (rows && rows_b ? 1 : 0) << ", \"gen_group\": " << gen_group
       << ", \"loop_groups\": " << (rows ? rest.ngroups : params.ngroups)
       << ", \"row_lds\": " << rlds_bytes << "}";
    return os.str();
}

}  // namespace qk
