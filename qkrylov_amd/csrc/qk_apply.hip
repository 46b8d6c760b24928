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
// Tiles are dealt to the 8 XCDs in 8 contiguous chunks (workgroup i runs on
// XCD i % 8), so each XCD's L2 holds the rows its waves are reading.
#include "qk_hamiltonian.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>

namespace qk {

namespace {

constexpr int kWaves = 4;            // waves per workgroup
constexpr int kBlock = 64 * kWaves;  // threads per workgroup
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
    for (u32 k = 0; k < g.nsup; ++k) c |= u32((s >> g.pos[k]) & 1) << k;
    return c;
}

// Tile index range of this wave.  Workgroup b lives on XCD b % 8; XCD k owns
// tiles [k*chunk, (k+1)*chunk) and its workgroups sweep that chunk together.
__device__ __forceinline__ void tile_range(const ApplyParams& p, u32 wave, u32& t0, u32& tend,
                                           u32& stride) {
    const u32 nb = gridDim.x;
    const u32 xcd = blockIdx.x % kXcd;
    const u32 kx = blockIdx.x / kXcd;
    const u32 per_xcd = nb / kXcd;
    const u32 chunk = p.tiles_per_xcd;
    const u32 start = xcd * chunk;
    t0 = start + kx * kWaves + wave;
    tend = min(start + chunk, p.ntiles);
    stride = per_xcd * kWaves;
}

template <typename S, int B>
__global__ void __launch_bounds__(kBlock)
k_apply(const ApplyParams p, const cplx* __restrict__ x, cplx* __restrict__ y, const u64 ld) {
    extern __shared__ __align__(16) unsigned char smem[];
    cplx* s_tab = reinterpret_cast<cplx*>(smem);
    u16* s_jlo = reinterpret_cast<u16*>(smem + size_t(p.ntab) * sizeof(cplx));

    for (u32 i = threadIdx.x; i < p.ntab; i += kBlock) s_tab[i] = p.tables[i];
    if ((p.nlo & 7) == 0) {
        const uint4* src = reinterpret_cast<const uint4*>(p.jlo);
        uint4* dst = reinterpret_cast<uint4*>(s_jlo);
        for (u32 i = threadIdx.x; i < p.nlo / 8; i += kBlock) dst[i] = src[i];
    } else {
        for (u32 i = threadIdx.x; i < p.nlo; i += kBlock) s_jlo[i] = p.jlo[i];
    }
    __syncthreads();

    const u32 wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const u32 lane = threadIdx.x & 63;
    const S lomask = (S(1) << p.L) - 1;
    u32 t, tend, stride;
    tile_range(p, wave, t, tend, stride);

    for (; t < tend; t += stride) {
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

        cplx acc[B];
#pragma unroll
        for (int v = 0; v < B; ++v) acc[v] = cplx{0.0, 0.0};
        cplx dsum{0.0, 0.0};

        u32 g = 0;
        // diagonal groups
        for (; g < p.n_diag; ++g) {
            const DevGroup G = p.groups[g];
            cplx amp = s_tab[G.tab + config_of(s, G)];
            if (G.jw && (popc_s(s & S(G.jw)) & 1)) { amp.re = -amp.re; amp.im = -amp.im; }
            dsum.re += amp.re;
            dsum.im += amp.im;
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
            if (__popc(tlo & u32(p.m1lo)) * p.ncls2 + __popc(tlo & u32(p.m2lo)) != cls) continue;
            if (G.jw && (popc_s(s & S(G.jw)) & 1)) { amp.re = -amp.re; amp.im = -amp.im; }
            const u64 b = base + rl;
#pragma unroll
            for (int v = 0; v < B; ++v) cfma(acc[v], amp, x[v * ld + b]);
        }
        // high-half groups: another row, same column
        const u32 g_hi_end = g_lo_end + p.n_hi;
        for (; g < g_hi_end; ++g) {
            const DevGroup G = p.groups[g];
            const u32 hi2 = hi ^ u32(G.flip >> p.L);
            const u64 base2 = p.jhi[hi2];
            if (base2 == ~u64(0) || p.hi_cls[hi2] != cls) continue;   // uniform
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
            const u64 base2 = p.jhi[hi2];
            if (base2 == ~u64(0)) continue;                               // uniform
            const int cls2 = p.hi_cls[hi2];
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
#pragma unroll
        for (int v = 0; v < B; ++v) {
            cplx o = acc[v];
            if (p.n_diag) cfma(o, dsum, x[v * ld + a]);
            y[v * ld + a] = o;
        }
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
    u32 t, tend, stride;
    tile_range(p, wave, t, tend, stride);
    for (; t < tend; t += stride) {
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
    u32 t, tend, stride;
    tile_range(p, wave, t, tend, stride);
    for (; t < tend; t += stride) {
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
    hipLaunchKernelGGL((k_apply<S, B>), dim3(h.grid), dim3(kBlock), h.lds_bytes, st, h.params, x,
                       y, ld);
}

template <typename S>
void launch_apply_s(const Hamiltonian& h, const cplx* x, cplx* y, int nvec, u64 ld,
                    hipStream_t st) {
    int v = 0;
    while (v < nvec) {
        const int left = nvec - v;
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
        for (size_t k = 0; k < g.pos.size(); ++k) d.pos[k] = static_cast<unsigned char>(g.pos[k]);
        tables.insert(tables.end(), g.table.begin(), g.table.end());
        out.push_back(d);
    }
    return out;
}

}  // namespace

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

    // Tiles: 64 consecutive ranks of one row.
    std::vector<uint2> tiles;
    for (u32 hi : B.valid_hi()) {
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

    // Persistent grid: enough workgroups to fill every CU a few times over,
    // a multiple of the XCD count, never more than there are tiles to run.
    const int cus = device_cu_count();
    const int want_blocks = int((tiles.size() + kWaves - 1) / kWaves);
    int blocks = std::min(want_blocks, cus * 8);
    blocks = std::max(kXcd, (blocks + kXcd - 1) / kXcd * kXcd);
    grid = blocks;
    p.tiles_per_xcd = static_cast<u32>((tiles.size() + kXcd - 1) / kXcd);

    lds_bytes = int(p.ntab * sizeof(cplx) + ((p.nlo * sizeof(u16) + 15) / 16) * 16);
    if (lds_bytes > 160 * 1024) fail("group tables exceed LDS");
    if (lds_bytes > 64 * 1024) {
        auto set = [&](const void* f) {
            QK_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes));
        };
        if (state32) {
            set((const void*)k_apply<u32, 1>); set((const void*)k_apply<u32, 2>);
            set((const void*)k_apply<u32, 4>); set((const void*)k_apply<u32, 8>);
        } else {
            set((const void*)k_apply<u64, 1>); set((const void*)k_apply<u64, 2>);
            set((const void*)k_apply<u64, 4>); set((const void*)k_apply<u64, 8>);
        }
    }

    dparams = p;
    dparams.groups = d_dgroups.get<DevGroup>();
    dparams.tables = d_dtables.get<cplx>();
    dparams.ngroups = static_cast<u32>(dgroups.size());
    dparams.ntab = static_cast<u32>(dtables.size());
}

Hamiltonian::~Hamiltonian() {
    if (stream) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamDestroy(stream);
    }
}

void Hamiltonian::apply(const cplx* x, cplx* y, int nvec, u64 ld, hipStream_t st) const {
    if (dim() == 0 || nvec <= 0) return;
    if (!st) st = stream;
    if (state32) launch_apply_s<u32>(*this, x, y, nvec, ld, st);
    else launch_apply_s<u64>(*this, x, y, nvec, ld, st);
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
       << ", \"max_support\": " << prog.max_support << "}";
    return os.str();
}

}  // namespace qk
