// Matrix-free Hamiltonian on one MI355X: the compiled group program and the
// basis ranking tables resident in HBM, applied by hand-written HIP kernels.
// Replaces the reference's MatrixFreeHamiltonian (include/qkrylov/
// hamiltonian/matrix_free_hamiltonian.hpp:24-77), which compiles a CSR on the
// host; nothing here is proportional to the number of matrix entries.
#pragma once

#include "qk_device.hpp"
#include "qk_model.hpp"

#include <memory>
#include <vector>

namespace qk {

// Group descriptor as the kernels read it (scalar loads, 48 B).  Support
// positions are packed 6 bits each into `posw` so that no array is indexed
// at run time (a dynamically indexed member would live in scratch memory).
struct DevGroup {
    u64 flip;
    u64 jw;
    u64 posw;      // support bit k at bits [6k, 6k+6)
    u32 tab;       // offset of the first table entry (cplx units)
    u32 nsup;      // number of support bits
    u32 kind;      // GroupKind
    u32 hi_only;   // support and jw inside the high half
    u64 pad;
};
static_assert(sizeof(DevGroup) == 48, "DevGroup layout");

// Exchange fast path: every off-diagonal group flips two bits i, j with real
// amplitude a01 (bit i = 0, bit j = 1 before the flip) or a10, on a basis
// whose sectors count all bits (spin-1/2 Sz, spinless fermion N) or none.
struct XBond {
    u32 mlo;       // flip mask inside the low half
    u32 mhi;       // flip mask inside the high half (shifted down by L)
    u32 bi, bj;    // bit positions of the two flipped bits (bi < bj)
    double a01, a10;
};
static_assert(sizeof(XBond) == 32, "XBond layout");

// Self-contained tile for the exchange kernel: 64 consecutive ranks of one
// row, everything the wave needs in one 32-byte scalar load.
struct XTile {
    u64 base;      // jhi[hi]: global index of rank 0 of the row
    u32 pos0;      // lo_list position of lane 0 (class offset + r0)
    u32 hi;
    u32 r0;        // rank of lane 0 inside the row
    u32 nvalid;    // lanes holding a state (<= 64)
    double dh;     // diagonal part from high-half-only terms
};
static_assert(sizeof(XTile) == 32, "XTile layout");

// Active high-half bond of one row for the single-vector exchange kernel:
// the partner row's offset jhi[hi ^ mhi] and the amplitude for this row's
// bit pattern.  Rows keep their active bonds in bond order, so a wave reads
// them with independent scalar loads rather than a bond -> row-offset chain.
struct XRowNb {
    u64 base;
    double amp;
};
static_assert(sizeof(XRowNb) == 16, "XRowNb layout");

struct XParams {
    const XTile* xtiles;
    const u16* jlo;
    const u64* jhi;
    const int* hi_cls;
    const u32* cls_off;
    const u32* cls_cnt;
    const u32* lo_list;
    const double* dlo;      // diagonal part from low-half-only terms, lo_list order
    const double* dhi;      // diagonal part from high-half-only terms, per hi
    const uint2* tiles;
    const XBond* bonds;     // lo bonds, then hi bonds, then mixed bonds
    const DevGroup* dmix;   // diagonal groups spanning both halves
    const cplx* dmix_tab;
    u32 ntiles, n_lo, n_hi, n_mix, n_dmix, ndmix_tab;
    u32 nlo;
    int L;
    u32 tiles_per_xcd;
    // single-vector kernel only: when set, workgroup b also writes
    // dot_part[b] = sum over its states of Re(conj(x_a) y_a)
    double* dot_part;
    // single-vector kernel only: per-row bond records
    const XRowNb* hnb;      // row hi: hnb[hi * n_hi + e], e < hcnt[hi]
    const u32* hcnt;        // active high-half bonds per row
    const u64* mnb;         // mnb[hi * n_mix + m] = jhi[hi ^ mhi] of mixed bond m
    const u32* vrow;        // null: records indexed by hi; else by vrow[hi]
};

// Tile of the multi-sector exchange kernel: 64 ranks of one row of sector
// `sec`, applied to all nvec vectors of that sector (vector v at
// xoff + v*ld of the concatenated buffers).
struct MTile {
    u64 base;      // jhi[hi] of the sector
    u64 xoff;      // offset of the sector's first vector
    u64 ld;        // sector dimension (vector stride)
    u32 pos0, hi, r0, nvalid;
    u32 sec, nvec;
    double dh;
};
static_assert(sizeof(MTile) == 56, "MTile layout");

struct MParams {
    const MTile* tiles;
    const u64* const* jhis;   // per-sector row offsets
    u32 ntiles, tiles_per_xcd;
};

// Bond records of one sector's rows for the single-vector exchange kernels
// (the per-row fields of XParams).
struct XRecs {
    const XRowNb* hnb;
    const u32* hcnt;
    const u64* mnb;
    const u32* vrow;
};

// Multi-sector single-vector kernel: the tiles of every sector in one array
// (sector s from xt[t0]) and each sector's records, a unit (one vector of
// one sector, at xoff of the concatenated buffers; its tile partial sums at
// part[poff + tile - t0]), and the tile list: entry (tile in xt, unit,
// sector), an empty tile (nvalid 0) for padding; the sector rides in the
// entry so that the tile, unit and sector loads go out together.
struct MsSec {
    XRecs rec;
    u32 t0, pad;
};
struct MsUnit {
    u64 xoff;
    u32 sec, poff;
};
static_assert(sizeof(MsUnit) == 16, "MsUnit layout");
struct MsParams {
    const uint4* tiles;
    const XTile* xt;
    const MsUnit* units;
    const MsSec* secs;
    double* part;   // null, or per tile Re<x, y> over its states
    u32 ntiles, tiles_per_xcd;
};

// Everything one apply launch needs, passed by value as the kernel argument.
struct ApplyParams {
    const u16* jlo;
    const u64* jhi;
    const int* hi_cls;
    const u32* cls_off;
    const u32* cls_cnt;
    const u32* lo_list;
    const uint2* tiles;
    const DevGroup* groups;
    const cplx* tables;
    u64 m1lo, m2lo;
    u32 ntiles;
    u32 ngroups;
    u32 n_diag, n_lo, n_hi, n_mix;
    u32 ntab;       // table entries copied to LDS
    u32 nlo;        // 2^L
    int L;
    int ncls2;
    u32 tiles_per_xcd;
};

// Row-program form of the groups that lie inside one half of the state
// (one basis): low-half groups whose support and
// Jordan-Wigner bits are in the low half act inside a row the same way for
// every row of a class, so per class they become a padded column-major table
// of (target rank, amplitude id) per rank; high-half groups inside the high
// half act the same way on every state of a row, so per row they become a
// list of (partner row offset, amplitude); diagonal groups inside one half
// become per-state (low) and per-row (high) sums.  Every other group stays
// in the kernel's group loop.
struct RowEntry {
    u64 base;
    cplx amp;
};
static_assert(sizeof(RowEntry) == 24, "RowEntry layout");

// A group across the split, seen from one row: the row fixes the group's
// high-half bits, so only its low-half sub-table (offset `tab` in the group
// table), its target row and the parity of its high-half Jordan-Wigner bits
// remain.  desc_sign = low-half descriptor index << 1 | that parity.
struct RowMix {
    u64 base;     // jhi of the target row
    u32 tab;      // sub-table over the low-half support configuration
    u32 desc_sign;
    int cls;      // class of the target row
    u32 pad;
};
static_assert(sizeof(RowMix) == 24, "RowMix layout");

struct RowParams {
    const cplx* dlo;        // lo_list order: diagonal groups inside the low half
    const cplx* dhi;        // per row (slot as rptr): diagonal groups inside the high half
    const u32* ell;         // class c, rank r, entry k: ell[ell_off[c] + k * cnt_c + r]
    const u32* ell_off;     //   entry = target rank | amplitude id << 16, ~0u: none
    const u32* ell_w;       // entries per rank of class c
    const cplx* amps;       // amplitude ids -> values (copied to LDS)
    const u32* rptr;        // row hi: rent[rptr[v] .. rptr[v + 1]), v = hi or vrow[hi]
    const RowEntry* rent;
    const u32* vrow;        // null: dhi and rptr indexed by hi; else by vrow[hi]
    const u32* mptr;        // groups across the split, row slot v: ment[mptr[v] .. mptr[v + 1])
    const RowMix* ment;     //   (null: none)
    const DevGroup* mdesc;  // low-half descriptors: flip, jw, support of the low half
    u32 nmdesc;             //   (copied to LDS)
    u32 namp;
    u32 ndiag;              // diagonal groups folded into dlo and dhi
};

// Row tables of the basis the targets are ranked in, when it differs from
// the basis whose rows are enumerated (Operator); unused otherwise.
struct ColTables {
    const u64* jhi = nullptr;
    const int* hi_cls = nullptr;
    u64 ld_out = 0;
};

struct MlpStates;   // qk_mlp.hpp

struct Hamiltonian {
    Hamiltonian(std::shared_ptr<const Basis> basis, SiteKind site, const OpSum& ops);
    ~Hamiltonian();
    Hamiltonian(const Hamiltonian&) = delete;
    Hamiltonian& operator=(const Hamiltonian&) = delete;

    u64 dim() const { return basis->size(); }

    // y[v*ld + a] = sum_b H^(gather)[a,b] x[v*ld + b], v < nvec, device pointers.
    void apply(const cplx* x, cplx* y, int nvec, u64 ld, hipStream_t st) const;
    // One vector on the single-vector exchange kernel, also writing
    // dot_parts() partial sums of Re<x, y> into part; false (nothing launched)
    // when that kernel is not this Hamiltonian's single-vector path.
    bool apply_dot(const cplx* x, cplx* y, double* part, hipStream_t st) const;
    int dot_parts() const { return xgrid_s; }
    // apply_dot is available and apply() of nvec vectors runs that same
    // kernel vector by vector
    bool fused_dot(int nvec) const {
        return xchg && xs_fn && dim() > 0 && (vec_group == 1 || nvec == 1);
    }
    // Full complex diagonal in OpSum order (device output, dim entries).
    void diagonal(cplx* d, hipStream_t st) const;
    // phi[k*ld + a] = N^-1/2 sum_j e^{i 2 pi k j / N} Sz_j psi[a], k < nq.
    void sz_q(const cplx* psi, cplx* phi, int nq, u64 ld, hipStream_t st) const;

    std::string info() const;
    // Time the kernel variants (exchange path: states per lane, bond batch,
    // vectors per launch; otherwise: row-program or group-loop kernel for
    // single vectors and blocks, vectors per launch) for blocks of nvec
    // vectors and keep the fastest; returns the timings (JSON).
    std::string autotune(int nvec);
    void setup_exchange();
    void setup_rows();
    // waves: 8 or 16 per workgroup; 0 chooses by size (QKRYLOV_XS_WAVES)
    void set_single_kernel(int states_per_lane, int bond_batch, int waves = 0);
    std::vector<XTile> xtiles_for(u32 span) const;

    std::shared_ptr<const Basis> basis;
    SiteKind site;
    OpSum ops;
    Program prog;
    hipStream_t stream = nullptr;
    int device = 0;

    // device-resident tables
    DevBuf d_jlo, d_jhi, d_hicls, d_clsoff, d_clscnt, d_lolist, d_tiles, d_groups, d_tables;
    DevBuf d_dgroups, d_dtables;   // per-term diagonal program
    DevBuf d_xbonds, d_dlo, d_dhi, d_dmix, d_dmixtab, d_xtiles, d_xtiles_s;   // exchange path
    DevBuf d_hnb, d_hcnt, d_mnb, d_vrow;   // per-row bond records (single-vector kernel)
    XParams xparams_s{};        // single-vector kernel, spl states per lane
    int xgrid_s = 0;
    int spl = 0;                   // states per lane of the single-vector kernel
    int bond_batch = 1;            // bonds whose loads one wave issues together
    int vec_group = 2;             // vectors per block-kernel launch (exchange path)
    const void* xs_fn = nullptr;   // the single-vector kernel for (spl, bond_batch)
    bool xsym = false;             // every low-half and straddling bond has a01 == a10
    std::vector<double> dhi_host;  // high-half diagonal per row
    std::vector<double> dlo_host;  // low-half diagonal, lo_list order
    std::vector<XBond> bonds_host;
    std::vector<DevGroup> dmix_host;   // diagonal groups across the split
    std::vector<cplx> dmixtab_host;
    std::vector<uint2> tiles_host;   // (hi, r0) per tile, launch order
    ApplyParams params{};
    ApplyParams dparams{};
    XParams xparams{};
    bool xchg = false;          // exchange fast path selected
    int xlds_bytes = 0;
    int xgrid = 0;
    bool force_generic = false; // QKRYLOV_APPLY=generic
    // row-program path (setup_rows): the remaining groups, their tables and
    // the row-program tables
    bool rows = false;          // row-program tables built
    bool rows_1 = true;         // single-vector launches use them
    bool rows_b = true;         // block launches use them
    int gen_group = 8;          // vectors per block launch off the exchange path
    ApplyParams rest{};
    RowParams rparams{};
    int rlds_bytes = 0;
    DevBuf d_rgroups, d_rtables, d_rdlo, d_rdhi, d_ell, d_elloff, d_ellw, d_amps, d_rptr, d_rent, d_rvrow;
    DevBuf d_mptr, d_ment, d_mdesc;
    bool state32 = true;
    int lds_bytes = 0;
    int dlds_bytes = 0;
    int grid = 0;
    int xs_res = 0;   // workgroups of the single-vector exchange kernel resident per CU
    int xs_wv = 16;   // waves per workgroup of xs_fn

    // scratch for host-buffer entry points
    mutable DevBuf scratch_x, scratch_y;
    // per-state inputs of the learned Davidson preconditioner, built on first
    // use (mlp_states_of) and kept for the Hamiltonian's lifetime
    mutable std::shared_ptr<MlpStates> mlp_states;
};

// y = H_s x for several Hamiltonians (sectors of one model) and several
// vectors each, in one launch per call: the sectors must be on the exchange
// path with identical bonds and diagonal tables (same OpSum, site and number
// of sites; only the sector differs).  Vectors of group g live at
// off_g + v*dim_g of the buffers passed to launch().
class MultiApply {
public:
    struct Group {
        const Hamiltonian* h;
        int nvec;
        u64 off;
    };
    // false when the groups cannot share one launch (callers then apply per group)
    bool build(const std::vector<Group>& groups, hipStream_t st);
    // part (tile_parts() only): unit u's tile t writes Re<x, y> over the
    // tile's states to part[u * part_stride() + t], t < unit_tiles()[u];
    // units are the groups' vectors in order
    void launch(const cplx* x, cplx* y, hipStream_t st, double* part = nullptr) const;
    bool ok() const { return ok_; }
    bool tile_parts() const { return ok_ && units_ && pstride_ > 0; }
    u32 part_stride() const { return pstride_; }
    const std::vector<u32>& unit_tiles() const { return utiles_; }

private:
    // per-vector units on XCD-resident tiles (k_apply_xchg_ms); false: the
    // block kernel k_apply_xchg_multi
    bool build_units(const std::vector<const Group*>& gs, hipStream_t st);
    bool ok_ = false;
    bool units_ = false;
    const Hamiltonian* h0_ = nullptr;
    MParams mp_{};
    DevBuf tiles_, jhis_;
    int grid_ = 0;
    int vb_ = 8;
    MsParams msp_{};
    DevBuf ms_tiles_, ms_units_, ms_secs_, ms_xt_;
    const void* ms_fn_ = nullptr;
    int ms_wv_ = 16;   // waves per workgroup of ms_fn_
    u32 pstride_ = 0;
    std::vector<u32> utiles_;
};

// y = O x for an OpSum O taking states of basis `in` to basis `out`: both of
// the same site kind and size and with the same counted masks (e.g. two Sz
// sectors, two particle-number sectors), so the low-half ranks agree; O need
// not be Hermitian nor conserve the sector (S+(q), c_k).  The gather kernel
// computes sum_b conj(amp(a -> b)) x[b] over the rows a of `out`, i.e. the
// conjugate transpose of what it compiles, so it compiles O's adjoint:
// coefficients conjugated, factor order reversed, each factor replaced by
// its adjoint (Sp <-> Sm, C <-> Cdag, CUp <-> CdagUp, CDn <-> CdagDn; Sz, Sx,
// Sy, N, Nup, Ndn, Nupdn, Id are self-adjoint).  No reference counterpart
// (the reference applies only Hamiltonians within one basis).
OpSum adjoint(const OpSum& ops);

class Operator {
public:
    Operator(std::shared_ptr<const Basis> in, std::shared_ptr<const Basis> out, SiteKind site,
             const OpSum& ops);
    u64 dim_in() const { return in_->size(); }
    u64 dim_out() const { return rows_->dim(); }
    // y[v*ld_out + b] = sum_a <b|O|a> x[v*ld_in + a], v < nvec (device pointers)
    void apply(const cplx* x, cplx* y, int nvec, u64 ld_in, u64 ld_out, hipStream_t st) const;
    hipStream_t stream() const { return rows_->stream; }

private:
    std::shared_ptr<const Basis> in_;
    std::unique_ptr<Hamiltonian> rows_;   // out's rows, O's adjoint compiled
    DevBuf jhi_in_, hicls_in_;
};

}  // namespace qk
