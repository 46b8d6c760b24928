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

// Group descriptor as the kernels read it (scalar loads, 48 B).
struct DevGroup {
    u64 flip;
    u64 jw;
    u32 tab;       // offset of the first table entry (cplx units)
    u32 nsup;      // number of support bits
    u32 kind;      // GroupKind
    u32 hi_only;   // support and jw inside the high half
    unsigned char pos[8];
    u64 pad;
};
static_assert(sizeof(DevGroup) == 48, "DevGroup layout");

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

struct Hamiltonian {
    Hamiltonian(std::shared_ptr<const Basis> basis, SiteKind site, const OpSum& ops);
    ~Hamiltonian();
    Hamiltonian(const Hamiltonian&) = delete;
    Hamiltonian& operator=(const Hamiltonian&) = delete;

    u64 dim() const { return basis->size(); }

    // y[v*ld + a] = sum_b H^(gather)[a,b] x[v*ld + b], v < nvec, device pointers.
    void apply(const cplx* x, cplx* y, int nvec, u64 ld, hipStream_t st) const;
    // Full complex diagonal in OpSum order (device output, dim entries).
    void diagonal(cplx* d, hipStream_t st) const;
    // phi[k*ld + a] = N^-1/2 sum_j e^{i 2 pi k j / N} Sz_j psi[a], k < nq.
    void sz_q(const cplx* psi, cplx* phi, int nq, u64 ld, hipStream_t st) const;

    std::string info() const;

    std::shared_ptr<const Basis> basis;
    SiteKind site;
    OpSum ops;
    Program prog;
    hipStream_t stream = nullptr;
    int device = 0;

    // device-resident tables
    DevBuf d_jlo, d_jhi, d_hicls, d_clsoff, d_clscnt, d_lolist, d_tiles, d_groups, d_tables;
    DevBuf d_dgroups, d_dtables;   // per-term diagonal program
    ApplyParams params{};
    ApplyParams dparams{};
    bool state32 = true;
    int lds_bytes = 0;
    int dlds_bytes = 0;
    int grid = 0;

    // scratch for host-buffer entry points
    mutable DevBuf scratch_x, scratch_y;
};

}  // namespace qk
