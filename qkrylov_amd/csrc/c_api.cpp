// extern "C" boundary (include/qkrylov/c_api.h).  Same handle ownership,
// NULL checks and status codes as the reference (src/c_api.cpp:33-626):
// creators return NULL on any failure, destroy(NULL) is a no-op, a NULL
// handle or buffer gives QKRYLOV_ERROR_INVALID_ARG, and every exception is
// caught here and returned as QKRYLOV_ERROR_EXCEPTION (its message is kept
// for qkrylov_last_error()).
#include "qkrylov/c_api.h"

#include "qk_hamiltonian.hpp"
#include "qk_hostla.hpp"
#include "qk_learned.hpp"
#include "qk_rng.hpp"
#include "qk_solvers.hpp"
#include "qk_store.hpp"
#include "qk_vec.hpp"

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace qk;

struct qkrylov_sector_t {
    Sector sector;
};
struct qkrylov_basis_t {
    std::shared_ptr<Basis> ptr;
};
struct qkrylov_site_t {
    SiteKind kind;
};
struct qkrylov_opsum_t {
    OpSum opsum;
};
struct qkrylov_hamiltonian_t {
    std::unique_ptr<Hamiltonian> ptr;
};
struct qkrylov_mlp_t {
    std::unique_ptr<Mlp> ptr;
};
struct qkrylov_operator_t {
    std::unique_ptr<Operator> ptr;
};

namespace {

thread_local std::string g_last_error;

void set_error(const char* msg) { g_last_error = msg ? msg : ""; }

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return QKRYLOV_SUCCESS;
    } catch (const std::exception& e) {
        set_error(e.what());
        return QKRYLOV_ERROR_EXCEPTION;
    } catch (...) {
        set_error("unknown exception");
        return QKRYLOV_ERROR_EXCEPTION;
    }
}

template <typename T, typename F>
T* guarded_create(F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        set_error(e.what());
        return nullptr;
    } catch (...) {
        set_error("unknown exception");
        return nullptr;
    }
}

qkrylov_basis_h make_basis(SiteKind k, int n, qkrylov_sector_h sector) {
    return guarded_create<qkrylov_basis_t>([&] {
        auto h = std::make_unique<qkrylov_basis_t>();
        const Sector sec = sector ? sector->sector : Sector();
        h->ptr = std::make_shared<Basis>(k, n, sec);
        return h.release();
    });
}

qkrylov_site_h make_site(SiteKind k) {
    return guarded_create<qkrylov_site_t>([&] {
        auto h = std::make_unique<qkrylov_site_t>();
        h->kind = k;
        return h.release();
    });
}

bool hok(qkrylov_hamiltonian_h h) { return h && h->ptr; }

// host complex FP64 -> device, apply, device -> host
void apply_host(const Hamiltonian& H, const cplx* x, cplx* y, int nvec) {
    const u64 d = H.dim();
    if (d == 0 || nvec <= 0) return;
    const size_t bytes = sizeof(cplx) * d * size_t(nvec);
    H.scratch_x.resize(bytes);
    H.scratch_y.resize(bytes);
    QK_HIP(hipMemcpyAsync(H.scratch_x.get(), x, bytes, hipMemcpyHostToDevice, H.stream));
    H.apply(H.scratch_x.get<cplx>(), H.scratch_y.get<cplx>(), nvec, d, H.stream);
    QK_HIP(hipMemcpyAsync(y, H.scratch_y.get(), bytes, hipMemcpyDeviceToHost, H.stream));
    QK_HIP(hipStreamSynchronize(H.stream));
}

std::vector<cplx> diag_host(const Hamiltonian& H) {
    const u64 d = H.dim();
    std::vector<cplx> out(d);
    if (d == 0) return out;
    DevBuf buf(sizeof(cplx) * d);
    H.diagonal(buf.get<cplx>(), H.stream);
    QK_HIP(hipMemcpyAsync(out.data(), buf.get(), sizeof(cplx) * d, hipMemcpyDeviceToHost, H.stream));
    QK_HIP(hipStreamSynchronize(H.stream));
    return out;
}

}  // namespace

namespace qk {
// the calling thread's message for qkrylov_last_error(), for entry points
// defined in other files
void capi_set_error(const char* msg) { set_error(msg); }
}  // namespace qk

extern "C" {

const char* qkrylov_last_error(void) { return g_last_error.c_str(); }

int qkrylov_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}
int qkrylov_device_info(char* buf, size_t len) {
    if (!buf || len == 0) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        int dev = 0;
        QK_HIP(hipGetDevice(&dev));
        hipDeviceProp_t p;
        QK_HIP(hipGetDeviceProperties(&p, dev));
        std::string s = std::string(p.gcnArchName) + " cus=" + std::to_string(p.multiProcessorCount) +
                        " mem_gb=" + std::to_string(p.totalGlobalMem >> 30) +
                        " l2_kb=" + std::to_string(p.l2CacheSize >> 10);
        std::strncpy(buf, s.c_str(), len - 1);
        buf[len - 1] = 0;
    });
}

/* ---------------- Sector ---------------- */
qkrylov_sector_h qkrylov_sector_create(void) {
    return guarded_create<qkrylov_sector_t>([] { return new qkrylov_sector_t(); });
}
void qkrylov_sector_destroy(qkrylov_sector_h sector) { delete sector; }
int qkrylov_sector_set_sz(qkrylov_sector_h s, int sz2) {
    if (!s) return QKRYLOV_ERROR_INVALID_ARG;
    s->sector.use_sz = true;
    s->sector.sz2 = sz2;
    return QKRYLOV_SUCCESS;
}
int qkrylov_sector_set_hubbard_particles(qkrylov_sector_h s, int nup, int ndn) {
    if (!s) return QKRYLOV_ERROR_INVALID_ARG;
    s->sector.use_nup = true;
    s->sector.use_ndn = true;
    s->sector.nup = nup;
    s->sector.ndn = ndn;
    return QKRYLOV_SUCCESS;
}
int qkrylov_sector_set_n(qkrylov_sector_h s, int n) {
    if (!s) return QKRYLOV_ERROR_INVALID_ARG;
    s->sector.use_n = true;
    s->sector.n = n;
    return QKRYLOV_SUCCESS;
}
int qkrylov_sector_set_nb(qkrylov_sector_h s, int nb) {
    if (!s) return QKRYLOV_ERROR_INVALID_ARG;
    s->sector.use_nb = true;
    s->sector.nb = nb;
    return QKRYLOV_SUCCESS;
}
int qkrylov_sector_get_sz(qkrylov_sector_h s, int* sz2_out, int* active_out) {
    if (!s) return QKRYLOV_ERROR_INVALID_ARG;
    if (active_out) *active_out = s->sector.use_sz ? 1 : 0;
    if (sz2_out) *sz2_out = s->sector.sz2;
    return QKRYLOV_SUCCESS;
}
int qkrylov_sector_get_hubbard_particles(qkrylov_sector_h s, int* nup_out, int* ndn_out,
                                         int* active_out) {
    if (!s) return QKRYLOV_ERROR_INVALID_ARG;
    if (active_out) *active_out = (s->sector.use_nup && s->sector.use_ndn) ? 1 : 0;
    if (nup_out) *nup_out = s->sector.nup;
    if (ndn_out) *ndn_out = s->sector.ndn;
    return QKRYLOV_SUCCESS;
}
int qkrylov_sector_get_n(qkrylov_sector_h s, int* n_out, int* active_out) {
    if (!s) return QKRYLOV_ERROR_INVALID_ARG;
    if (active_out) *active_out = s->sector.use_n ? 1 : 0;
    if (n_out) *n_out = s->sector.n;
    return QKRYLOV_SUCCESS;
}
int qkrylov_sector_get_nb(qkrylov_sector_h s, int* nb_out, int* active_out) {
    if (!s) return QKRYLOV_ERROR_INVALID_ARG;
    if (active_out) *active_out = s->sector.use_nb ? 1 : 0;
    if (nb_out) *nb_out = s->sector.nb;
    return QKRYLOV_SUCCESS;
}
int qkrylov_sector_reset(qkrylov_sector_h s) {
    if (!s) return QKRYLOV_ERROR_INVALID_ARG;
    s->sector = Sector();
    return QKRYLOV_SUCCESS;
}

/* ---------------- Basis ---------------- */
qkrylov_basis_h qkrylov_spinhalf_basis_create(int n, qkrylov_sector_h s) { return make_basis(SiteKind::SpinHalf, n, s); }
qkrylov_basis_h qkrylov_fermion_basis_create(int n, qkrylov_sector_h s) { return make_basis(SiteKind::Fermion, n, s); }
qkrylov_basis_h qkrylov_hubbard_basis_create(int n, qkrylov_sector_h s) { return make_basis(SiteKind::Hubbard, n, s); }
qkrylov_basis_h qkrylov_tj_basis_create(int n, qkrylov_sector_h s) { return make_basis(SiteKind::TJ, n, s); }
void qkrylov_basis_destroy(qkrylov_basis_h b) { delete b; }
uint64_t qkrylov_basis_dimension(qkrylov_basis_h b) {
    if (!b || !b->ptr) return 0;
    return b->ptr->size();
}
int qkrylov_basis_nsites(qkrylov_basis_h b) {
    if (!b || !b->ptr) return 0;
    return b->ptr->nsites();
}
uint64_t qkrylov_basis_state(qkrylov_basis_h b, uint64_t index) {
    if (!b || !b->ptr) return 0;
    try {
        return b->ptr->state(index);
    } catch (const std::exception& e) {
        set_error(e.what());
        return 0;
    }
}
int64_t qkrylov_basis_index(qkrylov_basis_h b, uint64_t s) {
    if (!b || !b->ptr) return -1;
    return b->ptr->index(s);
}
int qkrylov_basis_contains(qkrylov_basis_h b, uint64_t s) {
    if (!b || !b->ptr) return 0;
    return b->ptr->contains(s) ? 1 : 0;
}
int qkrylov_basis_states(qkrylov_basis_h b, uint64_t first, uint64_t count, uint64_t* out) {
    if (!b || !b->ptr || (!out && count)) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const Basis& B = *b->ptr;
        if (first + count > B.size() || first + count < first) fail("state range out of bounds");
        if (count == 0) return;
        // walk rows: one binary search, then sequential
        u64 i = first;
        u64 s0 = B.state(i);
        u64 hi = s0 >> B.L();
        size_t vh = std::lower_bound(B.valid_hi().begin(), B.valid_hi().end(), u32(hi)) - B.valid_hi().begin();
        for (u64 k = 0; k < count; ++k, ++i) {
            while (i >= B.jhi()[B.valid_hi()[vh]] + B.cls_cnt()[B.hi_cls()[B.valid_hi()[vh]]]) ++vh;
            const u32 h = B.valid_hi()[vh];
            const u64 r = i - B.jhi()[h];
            out[k] = (u64(h) << B.L()) | B.lo_list()[B.cls_off()[B.hi_cls()[h]] + r];
        }
    });
}
int qkrylov_basis_indices(qkrylov_basis_h b, const uint64_t* states, uint64_t count, int64_t* out) {
    if (!b || !b->ptr || ((!states || !out) && count)) return QKRYLOV_ERROR_INVALID_ARG;
    for (u64 k = 0; k < count; ++k) out[k] = b->ptr->index(states[k]);
    return QKRYLOV_SUCCESS;
}

/* ---------------- Site ---------------- */
qkrylov_site_h qkrylov_spinhalf_site_create(void) { return make_site(SiteKind::SpinHalf); }
qkrylov_site_h qkrylov_fermion_site_create(void) { return make_site(SiteKind::Fermion); }
qkrylov_site_h qkrylov_hubbard_site_create(void) { return make_site(SiteKind::Hubbard); }
qkrylov_site_h qkrylov_tj_site_create(void) { return make_site(SiteKind::TJ); }
void qkrylov_site_destroy(qkrylov_site_h s) { delete s; }

/* ---------------- OpSum ---------------- */
qkrylov_opsum_h qkrylov_opsum_create(void) {
    return guarded_create<qkrylov_opsum_t>([] { return new qkrylov_opsum_t(); });
}
void qkrylov_opsum_destroy(qkrylov_opsum_h o) { delete o; }
int qkrylov_opsum_clear(qkrylov_opsum_h o) {
    if (!o) return QKRYLOV_ERROR_INVALID_ARG;
    o->opsum.terms.clear();
    return QKRYLOV_SUCCESS;
}
int qkrylov_opsum_add_term_1body(qkrylov_opsum_h o, float cr, float ci, const char* op1, int s1) {
    if (!o || !op1) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        o->opsum.terms.push_back(OpTerm{cplx{double(cr), double(ci)}, {OpFactor{op1, s1}}});
    });
}
int qkrylov_opsum_add_term_2body(qkrylov_opsum_h o, float cr, float ci, const char* op1, int s1,
                                 const char* op2, int s2) {
    if (!o || !op1 || !op2) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        o->opsum.terms.push_back(
            OpTerm{cplx{double(cr), double(ci)}, {OpFactor{op1, s1}, OpFactor{op2, s2}}});
    });
}
int qkrylov_opsum_add_term_nbody(qkrylov_opsum_h o, float cr, float ci, int n, const char** ops,
                                 const int* sites) {
    if (!o || !ops || !sites || n <= 0) return QKRYLOV_ERROR_INVALID_ARG;
    for (int i = 0; i < n; ++i)
        if (!ops[i]) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        OpTerm t{cplx{double(cr), double(ci)}, {}};
        for (int i = 0; i < n; ++i) t.factors.push_back(OpFactor{ops[i], sites[i]});
        o->opsum.terms.push_back(std::move(t));
    });
}
int qkrylov_opsum_add_term_f64(qkrylov_opsum_h o, double cr, double ci, int n, const char** ops,
                               const int* sites) {
    if (!o || n < 0 || (n > 0 && (!ops || !sites))) return QKRYLOV_ERROR_INVALID_ARG;
    for (int i = 0; i < n; ++i)
        if (!ops[i]) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        OpTerm t{cplx{cr, ci}, {}};
        for (int i = 0; i < n; ++i) t.factors.push_back(OpFactor{ops[i], sites[i]});
        o->opsum.terms.push_back(std::move(t));
    });
}
size_t qkrylov_opsum_size(qkrylov_opsum_h o) { return o ? o->opsum.terms.size() : 0; }

/* ---------------- Hamiltonian ---------------- */
qkrylov_hamiltonian_h qkrylov_hamiltonian_create(qkrylov_basis_h basis, qkrylov_site_h site,
                                                qkrylov_opsum_h opsum) {
    if (!basis || !basis->ptr || !site || !opsum) return nullptr;
    return guarded_create<qkrylov_hamiltonian_t>([&] {
        auto h = std::make_unique<qkrylov_hamiltonian_t>();
        h->ptr = std::make_unique<Hamiltonian>(basis->ptr, site->kind, opsum->opsum);
        return h.release();
    });
}
void qkrylov_hamiltonian_destroy(qkrylov_hamiltonian_h h) { delete h; }
uint64_t qkrylov_hamiltonian_dimension(qkrylov_hamiltonian_h h) {
    if (!hok(h)) return 0;
    return h->ptr->dim();
}
int qkrylov_hamiltonian_info(qkrylov_hamiltonian_h h, char* buf, size_t len) {
    if (!hok(h) || !buf || len == 0) return QKRYLOV_ERROR_INVALID_ARG;
    const std::string s = h->ptr->info();
    std::strncpy(buf, s.c_str(), len - 1);
    buf[len - 1] = 0;
    return QKRYLOV_SUCCESS;
}
int qkrylov_hamiltonian_autotune(qkrylov_hamiltonian_h h, int nvec, char* buf, size_t len) {
    if (!hok(h) || nvec <= 0) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const std::string s = h->ptr->autotune(nvec);
        if (buf && len) {
            std::strncpy(buf, s.c_str(), len - 1);
            buf[len - 1] = 0;
        }
    });
}
int qkrylov_block_dot_f64(const double* V, int m, const double* W, int nw, uint64_t n, double* G) {
    if (!V || !W || !G || m <= 0 || nw <= 0 || m > 64 || nw > 16) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        hipStream_t st = nullptr;
        QK_HIP(hipStreamCreate(&st));
        struct Guard {
            hipStream_t s;
            ~Guard() { (void)hipStreamDestroy(s); }
        } guard{st};
        const u64 nn = std::max<u64>(n, 1);
        DevBuf dv(sizeof(cplx) * nn * m), dw(sizeof(cplx) * nn * nw), g(sizeof(cplx) * m * nw),
            pv(sizeof(void*) * m), pw(sizeof(void*) * nw), scratch;
        std::vector<const cplx*> hv(m), hw(nw);
        for (int i = 0; i < m; ++i) hv[i] = dv.get<cplx>() + u64(i) * n;
        for (int j = 0; j < nw; ++j) hw[j] = dw.get<cplx>() + u64(j) * n;
        if (n) {
            QK_HIP(hipMemcpyAsync(dv.get(), V, sizeof(cplx) * n * m, hipMemcpyHostToDevice, st));
            QK_HIP(hipMemcpyAsync(dw.get(), W, sizeof(cplx) * n * nw, hipMemcpyHostToDevice, st));
        }
        QK_HIP(hipMemcpyAsync(pv.get(), hv.data(), sizeof(void*) * m, hipMemcpyHostToDevice, st));
        QK_HIP(hipMemcpyAsync(pw.get(), hw.data(), sizeof(void*) * nw, hipMemcpyHostToDevice, st));
        vec::block_dot(pv.get<const cplx*>(), m, pw.get<const cplx*>(), nw, n, g.get<cplx>(), scratch, st);
        QK_HIP(hipMemcpyAsync(G, g.get(), sizeof(cplx) * m * nw, hipMemcpyDeviceToHost, st));
        QK_HIP(hipStreamSynchronize(st));
    });
}
int qkrylov_hamiltonian_sync(qkrylov_hamiltonian_h h) {
    if (!hok(h)) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] { QK_HIP(hipStreamSynchronize(h->ptr->stream)); });
}

int qkrylov_hamiltonian_apply(qkrylov_hamiltonian_h h, const float* xr, const float* xi, float* yr,
                              float* yi) {
    if (!hok(h) || !xr || !yr) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const u64 d = h->ptr->dim();
        std::vector<cplx> x(d), y(d);
        for (u64 i = 0; i < d; ++i) x[i] = cplx{double(xr[i]), xi ? double(xi[i]) : 0.0};
        apply_host(*h->ptr, x.data(), y.data(), 1);
        for (u64 i = 0; i < d; ++i) {
            yr[i] = float(y[i].re);
            if (yi) yi[i] = float(y[i].im);
        }
    });
}
int qkrylov_hamiltonian_apply_complex(qkrylov_hamiltonian_h h, const float* x, float* y) {
    if (!hok(h) || !x || !y) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const u64 d = h->ptr->dim();
        std::vector<cplx> xc(d), yc(d);
        for (u64 i = 0; i < d; ++i) xc[i] = cplx{double(x[2 * i]), double(x[2 * i + 1])};
        apply_host(*h->ptr, xc.data(), yc.data(), 1);
        for (u64 i = 0; i < d; ++i) {
            y[2 * i] = float(yc[i].re);
            y[2 * i + 1] = float(yc[i].im);
        }
    });
}
int qkrylov_hamiltonian_diagonal(qkrylov_hamiltonian_h h, float* out) {
    if (!hok(h) || !out) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const std::vector<cplx> d = diag_host(*h->ptr);
        for (size_t i = 0; i < d.size(); ++i) out[i] = float(d[i].re);
    });
}
int qkrylov_hamiltonian_apply_f64(qkrylov_hamiltonian_h h, const double* x, double* y, int nvec) {
    if (!hok(h) || !x || !y || nvec <= 0) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        apply_host(*h->ptr, reinterpret_cast<const cplx*>(x), reinterpret_cast<cplx*>(y), nvec);
    });
}
int qkrylov_hamiltonian_apply_device_f64(qkrylov_hamiltonian_h h, const void* x, void* y, int nvec,
                                         uint64_t ld, void* stream) {
    if (!hok(h) || !x || !y || nvec <= 0) return QKRYLOV_ERROR_INVALID_ARG;
    if (nvec > 1 && ld < h->ptr->dim()) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        h->ptr->apply(static_cast<const cplx*>(x), static_cast<cplx*>(y), nvec,
                      ld ? ld : h->ptr->dim(), static_cast<hipStream_t>(stream));
    });
}
qkrylov_operator_h qkrylov_operator_create(qkrylov_basis_h in, qkrylov_basis_h out, qkrylov_site_h site,
                                          qkrylov_opsum_h opsum) {
    if (!in || !in->ptr || !out || !out->ptr || !site || !opsum) return nullptr;
    return guarded_create<qkrylov_operator_t>([&] {
        auto h = std::make_unique<qkrylov_operator_t>();
        h->ptr = std::make_unique<Operator>(in->ptr, out->ptr, site->kind, opsum->opsum);
        return h.release();
    });
}
void qkrylov_operator_destroy(qkrylov_operator_h op) { delete op; }
uint64_t qkrylov_operator_dimension_in(qkrylov_operator_h op) { return op && op->ptr ? op->ptr->dim_in() : 0; }
uint64_t qkrylov_operator_dimension_out(qkrylov_operator_h op) { return op && op->ptr ? op->ptr->dim_out() : 0; }
int qkrylov_operator_apply_f64(qkrylov_operator_h op, const double* x, double* y, int nvec) {
    if (!op || !op->ptr || !y || nvec <= 0) return QKRYLOV_ERROR_INVALID_ARG;
    const Operator& O = *op->ptr;
    if (!x && O.dim_in() > 0) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const u64 di = O.dim_in(), dout = O.dim_out();
        if (dout == 0) return;
        hipStream_t st = O.stream();
        DevBuf xb(sizeof(cplx) * std::max<u64>(di, 1) * nvec), yb(sizeof(cplx) * dout * nvec);
        if (di) QK_HIP(hipMemcpyAsync(xb.get(), x, sizeof(cplx) * di * nvec, hipMemcpyHostToDevice, st));
        O.apply(xb.get<cplx>(), yb.get<cplx>(), nvec, std::max<u64>(di, 1), dout, st);
        QK_HIP(hipMemcpyAsync(y, yb.get(), sizeof(cplx) * dout * nvec, hipMemcpyDeviceToHost, st));
        QK_HIP(hipStreamSynchronize(st));
    });
}
int qkrylov_operator_apply_device_f64(qkrylov_operator_h op, const void* x, void* y, int nvec, uint64_t ld_in,
                                      uint64_t ld_out, void* stream) {
    if (!op || !op->ptr || !y || nvec <= 0) return QKRYLOV_ERROR_INVALID_ARG;
    const Operator& O = *op->ptr;
    if (!x && O.dim_in() > 0) return QKRYLOV_ERROR_INVALID_ARG;
    if (ld_in == 0) ld_in = O.dim_in();
    if (ld_out == 0) ld_out = O.dim_out();
    if (nvec > 1 && (ld_in < O.dim_in() || ld_out < O.dim_out())) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        O.apply(static_cast<const cplx*>(x), static_cast<cplx*>(y), nvec, ld_in, ld_out,
                static_cast<hipStream_t>(stream));
    });
}

int qkrylov_hamiltonian_diagonal_f64(qkrylov_hamiltonian_h h, double* out) {
    if (!hok(h) || !out) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const std::vector<cplx> d = diag_host(*h->ptr);
        std::memcpy(out, d.data(), sizeof(cplx) * d.size());
    });
}
int qkrylov_sz_q_device_f64(qkrylov_hamiltonian_h h, const void* psi, void* phi, int nq,
                            uint64_t ld, void* stream) {
    if (!hok(h) || !psi || !phi || nq <= 0) return QKRYLOV_ERROR_INVALID_ARG;
    if (ld < h->ptr->dim()) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        h->ptr->sz_q(static_cast<const cplx*>(psi), static_cast<cplx*>(phi), nq, ld,
                     static_cast<hipStream_t>(stream));
    });
}

/* ---------------- Solvers ---------------- */
int qkrylov_lanczos_ground_state(qkrylov_hamiltonian_h h, int maxiter, float tol,
                                 qkrylov_lanczos_result_c_t* result) {
    return qkrylov_lanczos_ground_state_complex(h, maxiter, tol, result, nullptr);
}
int qkrylov_lanczos_ground_state_complex(qkrylov_hamiltonian_h h, int maxiter, float tol,
                                         qkrylov_lanczos_result_c_t* result, float* eig) {
    if (!hok(h) || !result) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const LanczosResult r = lanczos_ground_state(*h->ptr, maxiter, double(tol), nullptr, eig != nullptr);
        result->energy = float(r.energy);
        result->iterations = r.iterations;
        result->converged = r.converged ? 1 : 0;
        if (eig && !r.eigenvector.empty())
            for (size_t i = 0; i < r.eigenvector.size(); ++i) {
                eig[2 * i] = float(r.eigenvector[i].re);
                eig[2 * i + 1] = float(r.eigenvector[i].im);
            }
    });
}
int qkrylov_lanczos_ground_state_f64(qkrylov_hamiltonian_h h, int maxiter, double tol,
                                     const double* start, qkrylov_lanczos_result_f64_t* result,
                                     double* eig) {
    if (!hok(h) || !result) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const LanczosResult r = lanczos_ground_state(*h->ptr, maxiter, tol,
                                                     reinterpret_cast<const cplx*>(start), eig != nullptr);
        result->energy = r.energy;
        result->iterations = r.iterations;
        result->converged = r.converged ? 1 : 0;
        result->second_passes = r.second_passes;
        if (eig && !r.eigenvector.empty())
            std::memcpy(eig, r.eigenvector.data(), sizeof(cplx) * r.eigenvector.size());
    });
}

int qkrylov_lanczos_checkpointed_f64(qkrylov_hamiltonian_h h, int maxiter, double tol,
                                     const double* start, const char* save_path, int save_every,
                                     int stop_after, const char* resume_path,
                                     qkrylov_lanczos_result_f64_t* result, double* eig) {
    if (!hok(h) || !result || save_every < 0 || stop_after < 0 || (stop_after > 0 && !save_path))
        return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        LanczosIO io;
        if (save_path) io.save_path = save_path;
        if (resume_path) io.resume_path = resume_path;
        io.save_every = save_every;
        io.stop_after = stop_after;
        const LanczosResult r = lanczos_run(*h->ptr, maxiter, tol, reinterpret_cast<const cplx*>(start),
                                            eig != nullptr, io);
        result->energy = r.energy;
        result->iterations = r.iterations;
        result->converged = r.converged ? 1 : 0;
        result->second_passes = r.second_passes;
        if (eig && !r.eigenvector.empty())
            std::memcpy(eig, r.eigenvector.data(), sizeof(cplx) * r.eigenvector.size());
    });
}

int qkrylov_eigenpairs_save_f64(const char* path, uint64_t dim, int n, const double* values,
                                const double* vectors_complex, uint64_t fingerprint) {
    if (!path || n < 0 || (n > 0 && !values)) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        save_eigenpairs(path, dim, fingerprint, n, values, reinterpret_cast<const cplx*>(vectors_complex));
    });
}

int qkrylov_eigenpairs_info(const char* path, uint64_t* dim, int* n, int* has_vectors,
                            uint64_t* fingerprint) {
    if (!path) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const EigenFile e = load_eigenpairs(path, false);
        if (dim) *dim = e.dim;
        if (n) *n = int(e.values.size());
        if (has_vectors) *has_vectors = e.has_vectors ? 1 : 0;
        if (fingerprint) *fingerprint = e.fingerprint;
    });
}

int qkrylov_eigenpairs_load_f64(const char* path, double* values_out, double* vectors_complex_out) {
    if (!path || !values_out) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const EigenFile e = load_eigenpairs(path, true);
        if (!e.values.empty()) std::memcpy(values_out, e.values.data(), sizeof(double) * e.values.size());
        if (vectors_complex_out) {
            if (!e.has_vectors) fail("eigenpairs: the file holds no vectors");
            std::memcpy(vectors_complex_out, e.vectors.data(), sizeof(cplx) * e.vectors.size());
        }
    });
}

uint64_t qkrylov_hamiltonian_fingerprint(qkrylov_hamiltonian_h h) {
    if (!hok(h)) return 0;
    u64 fp = 0;
    (void)guarded([&] { fp = hamiltonian_fingerprint(*h->ptr); });
    return fp;
}

int qkrylov_davidson_lowest_complex(qkrylov_hamiltonian_h h, int n_eig, int max_subspace, float tol,
                                    float* evals, float* evecs, qkrylov_davidson_result_c_t* info) {
    if (!hok(h) || !evals || n_eig <= 0) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const DavidsonResult r = davidson_lowest(*h->ptr, n_eig, max_subspace, double(tol), evecs != nullptr);
        const size_t k = std::min(size_t(n_eig), r.eigenvalues.size());
        for (size_t i = 0; i < k; ++i) evals[i] = float(r.eigenvalues[i]);
        if (info) {
            info->iterations = r.iterations;
            info->converged = r.converged ? 1 : 0;
        }
        if (evecs) {
            const u64 d = h->ptr->dim();
            for (size_t idx = 0; idx < k && idx < r.eigenvectors.size(); ++idx) {
                float* dst = evecs + idx * 2 * d;
                const auto& v = r.eigenvectors[idx];
                for (size_t i = 0; i < d && i < v.size(); ++i) {
                    dst[2 * i] = float(v[i].re);
                    dst[2 * i + 1] = float(v[i].im);
                }
            }
        }
    });
}
int qkrylov_block_lanczos_f64(qkrylov_hamiltonian_h h, int block, int nev, int maxiter, double tol,
                              double* evals, double* evecs, int* iterations, int* converged) {
    if (!hok(h) || !evals || block <= 0 || nev <= 0 || nev > block || maxiter <= 0)
        return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const BlockLanczosResult r = block_lanczos_lowest(*h->ptr, block, nev, maxiter, tol, evecs != nullptr);
        for (size_t i = 0; i < r.eigenvalues.size(); ++i) evals[i] = r.eigenvalues[i];
        if (iterations) *iterations = r.iterations;
        if (converged) *converged = r.converged ? 1 : 0;
        if (evecs) {
            const u64 d = h->ptr->dim();
            for (size_t k = 0; k < r.eigenvectors.size(); ++k)
                std::memcpy(evecs + k * 2 * d, r.eigenvectors[k].data(), sizeof(cplx) * d);
        }
    });
}
int qkrylov_davidson_lowest_f64(qkrylov_hamiltonian_h h, int n_eig, int max_subspace, double tol,
                                double* evals, double* evecs, qkrylov_davidson_result_c_t* info) {
    if (!hok(h) || !evals || n_eig <= 0) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const DavidsonResult r = davidson_lowest(*h->ptr, n_eig, max_subspace, tol, evecs != nullptr);
        const size_t k = std::min(size_t(n_eig), r.eigenvalues.size());
        for (size_t i = 0; i < k; ++i) evals[i] = r.eigenvalues[i];
        if (info) {
            info->iterations = r.iterations;
            info->converged = r.converged ? 1 : 0;
        }
        if (evecs) {
            const u64 d = h->ptr->dim();
            for (size_t idx = 0; idx < k && idx < r.eigenvectors.size(); ++idx)
                std::memcpy(evecs + idx * 2 * d, r.eigenvectors[idx].data(), sizeof(cplx) * d);
        }
    });
}

// ---- learned Davidson preconditioner (no reference counterpart) ----
qkrylov_mlp_h qkrylov_mlp_create(unsigned seed) {
    return guarded_create<qkrylov_mlp_t>([&] {
        auto m = std::make_unique<qkrylov_mlp_t>();
        m->ptr = std::make_unique<Mlp>(seed);
        return m.release();
    });
}
void qkrylov_mlp_destroy(qkrylov_mlp_h m) { delete m; }
int qkrylov_mlp_num_params(void) { return kMlpParams; }
int qkrylov_mlp_get_params(qkrylov_mlp_h m, float* out) {
    if (!m || !m->ptr || !out) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] { m->ptr->get_params(out); });
}
int qkrylov_mlp_set_params(qkrylov_mlp_h m, const float* in) {
    if (!m || !m->ptr || !in) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] { m->ptr->set_params(in); });
}
int qkrylov_mlp_eval(qkrylov_mlp_h m, const float* features, uint64_t n, const float* grad_out,
                     float* o_out, float* grad_params) {
    if (!m || !m->ptr || (n && (!features || !o_out))) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] { m->ptr->eval_host(features, n, grad_out, o_out, grad_params); });
}
int qkrylov_mlp_features(qkrylov_hamiltonian_h h, double theta, float* out) {
    if (!hok(h) || !out) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const Hamiltonian& H = *h->ptr;
        const std::shared_ptr<const MlpStates> sp = mlp_states_of(H);
        const MlpStates& s = *sp;
        const u64 d = H.dim();
        DevBuf X(sizeof(u16) * std::max<u64>(d, 1) * kMlpIn);
        mlp_features(s, theta, X.get<u16>(), H.stream);
        std::vector<u16> xb(d * kMlpIn);
        if (d) QK_HIP(hipMemcpyAsync(xb.data(), X.get(), sizeof(u16) * xb.size(), hipMemcpyDeviceToHost, H.stream));
        QK_HIP(hipStreamSynchronize(H.stream));
        for (size_t i = 0; i < xb.size(); ++i) out[i] = bf16_to_f32(xb[i]);
    });
}
int qkrylov_mlp_correction_f64(qkrylov_mlp_h m, qkrylov_hamiltonian_h h, const double* r,
                               double theta, double* t_out, double* z_out, double* w_out,
                               double* w2_out) {
    if (!m || !m->ptr || !hok(h) || !r || !t_out) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const Hamiltonian& H = *h->ptr;
        const u64 d = H.dim();
        if (!d) return;
        MlpPreconditioner pre(*m->ptr, H);
        DevBuf b(sizeof(cplx) * d * 2);
        cplx* rd = b.get<cplx>();
        cplx* td = rd + d;
        QK_HIP(hipMemcpyAsync(rd, r, sizeof(cplx) * d, hipMemcpyHostToDevice, H.stream));
        pre.apply(H, rd, nullptr, theta, nullptr, td, H.stream);
        QK_HIP(hipMemcpyAsync(t_out, td, sizeof(cplx) * d, hipMemcpyDeviceToHost, H.stream));
        if (z_out) QK_HIP(hipMemcpyAsync(z_out, pre.z(), sizeof(cplx) * d, hipMemcpyDeviceToHost, H.stream));
        if (w_out) QK_HIP(hipMemcpyAsync(w_out, pre.w(), sizeof(cplx) * d, hipMemcpyDeviceToHost, H.stream));
        if (w2_out) QK_HIP(hipMemcpyAsync(w2_out, pre.w2(), sizeof(cplx) * d, hipMemcpyDeviceToHost, H.stream));
        QK_HIP(hipStreamSynchronize(H.stream));
    });
}
int qkrylov_mlp_pair_loss_f64(qkrylov_hamiltonian_h h, int npairs, const float* o, const double* z,
                              const double* w, const double* w2, const double* b, const double* u,
                              const double* theta, float* grad_o, double* loss) {
    if (!hok(h) || !o || !z || !w || !w2 || !b || !u || !theta || !loss || npairs <= 0)
        return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const Hamiltonian& H = *h->ptr;
        hipStream_t st = H.stream;
        const u64 n = H.dim();
        const u64 tot = u64(npairs) * n;
        if (!n) {
            *loss = 0.0;
            return;
        }
        DevBuf vb(sizeof(cplx) * tot * 5), ob(sizeof(float) * tot * kMlpOut),
            gb(sizeof(float) * tot * kMlpOut), tb(sizeof(double) * size_t(npairs));
        cplx* zd = vb.get<cplx>();
        const double* src[5] = {z, w, w2, b, u};
        for (int q = 0; q < 5; ++q)
            QK_HIP(hipMemcpyAsync(zd + u64(q) * tot, src[q], sizeof(cplx) * tot, hipMemcpyHostToDevice, st));
        QK_HIP(hipMemcpyAsync(ob.get(), o, sizeof(float) * tot * kMlpOut, hipMemcpyHostToDevice, st));
        QK_HIP(hipMemcpyAsync(tb.get(), theta, sizeof(double) * size_t(npairs), hipMemcpyHostToDevice, st));
        MlpLossWork wk;
        *loss = mlp_pair_loss(H, zd, zd + tot, zd + 2 * tot, zd + 3 * tot, zd + 4 * tot, tb.get<double>(), npairs,
                              ob.get<float>(), grad_o ? gb.get<float>() : nullptr, wk, st);
        if (grad_o)
            QK_HIP(hipMemcpyAsync(grad_o, gb.get(), sizeof(float) * tot * kMlpOut, hipMemcpyDeviceToHost, st));
        QK_HIP(hipStreamSynchronize(st));
    });
}
int qkrylov_jd_correction_f64(qkrylov_hamiltonian_h h, const double* r, const double* u, double theta,
                              int max_iter, double rtol, double* t_out, double* rel_res) {
    if (!hok(h) || !r || !u || !t_out || max_iter <= 0) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const Hamiltonian& H = *h->ptr;
        const u64 d = H.dim();
        DevBuf b(sizeof(cplx) * std::max<u64>(d, 1) * 3);
        cplx* rd = b.get<cplx>();
        cplx* ud = rd + d;
        cplx* td = ud + d;
        QK_HIP(hipMemcpyAsync(rd, r, sizeof(cplx) * d, hipMemcpyHostToDevice, H.stream));
        QK_HIP(hipMemcpyAsync(ud, u, sizeof(cplx) * d, hipMemcpyHostToDevice, H.stream));
        const double rr = jd_correction(H, rd, ud, theta, td, max_iter, rtol, H.stream);
        QK_HIP(hipMemcpyAsync(t_out, td, sizeof(cplx) * d, hipMemcpyDeviceToHost, H.stream));
        QK_HIP(hipStreamSynchronize(H.stream));
        if (rel_res) *rel_res = rr;
    });
}
int qkrylov_mlp_train_f64(qkrylov_mlp_h m, qkrylov_hamiltonian_h hs, int n_eig, int max_subspace,
                          int max_iter, int max_pairs, int epochs, int batch_pairs, double lr,
                          double* loss_history, qkrylov_mlp_train_info_t* info) {
    if (!m || !m->ptr || !hok(hs) || n_eig <= 0 || max_pairs <= 0 || epochs < 0 || batch_pairs <= 0)
        return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        MlpTrainConfig cfg;
        cfg.n_eig = n_eig;
        cfg.max_subspace = max_subspace;
        cfg.max_iter = std::max(1, max_iter);
        cfg.max_pairs = max_pairs;
        cfg.epochs = epochs;
        cfg.batch_pairs = batch_pairs;
        cfg.lr = lr;
        std::vector<double> hist;
        const MlpTrainStats st = train_mlp(*m->ptr, *hs->ptr, cfg, &hist);
        if (loss_history)
            for (int e = 0; e <= epochs; ++e) loss_history[e] = e < int(hist.size()) ? hist[e] : 0.0;
        if (info) {
            info->pairs = st.pairs;
            info->epochs = st.epochs;
            info->loss_first = st.loss_first;
            info->loss_last = st.loss_last;
            info->solve_rel_res = st.solve_rel_res;
        }
    });
}
int qkrylov_davidson_lowest_ex_f64(qkrylov_hamiltonian_h h, int n_eig, int max_subspace, double tol,
                                   int max_iter, qkrylov_mlp_h mlp, double* evals, double* evecs,
                                   qkrylov_davidson_result_c_t* info) {
    if (!hok(h) || !evals || n_eig <= 0 || max_iter <= 0 || (mlp && !mlp->ptr))
        return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        std::unique_ptr<MlpPreconditioner> pre;
        if (mlp) pre = std::make_unique<MlpPreconditioner>(*mlp->ptr, *h->ptr);
        const DavidsonResult r =
            davidson_lowest(*h->ptr, n_eig, max_subspace, tol, evecs != nullptr, pre.get(), max_iter);
        const size_t k = std::min(size_t(n_eig), r.eigenvalues.size());
        for (size_t i = 0; i < k; ++i) evals[i] = r.eigenvalues[i];
        if (info) {
            info->iterations = r.iterations;
            info->converged = r.converged ? 1 : 0;
        }
        if (evecs) {
            const u64 d = h->ptr->dim();
            for (size_t idx = 0; idx < k && idx < r.eigenvectors.size(); ++idx)
                std::memcpy(evecs + idx * 2 * d, r.eigenvectors[idx].data(), sizeof(cplx) * d);
        }
    });
}

int qkrylov_continued_fraction_coeffs_complex(qkrylov_hamiltonian_h h, const float* phi0, int n_iter,
                                              float* alphas, float* betas, float* norm, int* count) {
    if (!hok(h) || !phi0 || !alphas || !betas || !norm || n_iter <= 0) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const u64 d = h->ptr->dim();
        std::vector<cplx> p(d);
        for (u64 i = 0; i < d; ++i) p[i] = cplx{double(phi0[2 * i]), double(phi0[2 * i + 1])};
        const std::vector<CFResult> r = continued_fraction_block(*h->ptr, p.data(), 1, n_iter);
        const CFResult cf = r.empty() ? CFResult{} : r[0];
        *norm = float(cf.norm_phi0);
        for (size_t i = 0; i < cf.alphas.size() && i < size_t(n_iter); ++i) alphas[i] = float(cf.alphas[i]);
        for (size_t i = 0; i < cf.betas.size() && i < size_t(n_iter); ++i) betas[i] = float(cf.betas[i]);
        if (count) *count = int(cf.alphas.size());
    });
}
int qkrylov_continued_fraction_coeffs_f64(qkrylov_hamiltonian_h h, const double* phi0, int n_iter,
                                          double* alphas, double* betas, double* norm, int* count) {
    return qkrylov_continued_fraction_block_f64(h, phi0, 1, n_iter, alphas, betas, norm, count);
}
int qkrylov_continued_fraction_block_f64(qkrylov_hamiltonian_h h, const double* phi0, int nblock,
                                         int n_iter, double* alphas, double* betas, double* norms,
                                         int* counts) {
    if (!hok(h) || !phi0 || !alphas || !betas || !norms || n_iter <= 0 || nblock <= 0)
        return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const std::vector<CFResult> r =
            continued_fraction_block(*h->ptr, reinterpret_cast<const cplx*>(phi0), nblock, n_iter);
        for (int b = 0; b < nblock; ++b) {
            const CFResult& cf = r[b];
            norms[b] = cf.norm_phi0;
            for (int i = 0; i < n_iter; ++i) {
                alphas[size_t(b) * n_iter + i] = i < int(cf.alphas.size()) ? cf.alphas[i] : 0.0;
                betas[size_t(b) * n_iter + i] = i < int(cf.betas.size()) ? cf.betas[i] : 0.0;
            }
            if (counts) counts[b] = int(cf.alphas.size());
        }
    });
}

int qkrylov_continued_fraction_block_device_f64(qkrylov_hamiltonian_h h, const void* phi0_dev,
                                                int nblock, int n_iter, double* alphas, double* betas,
                                                double* norms, int* counts) {
    if (!hok(h) || !phi0_dev || !alphas || !betas || !norms || n_iter <= 0 || nblock <= 0)
        return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const std::vector<CFResult> r = continued_fraction_block_dev(
            *h->ptr, static_cast<const cplx*>(phi0_dev), nblock, n_iter);
        for (int b = 0; b < nblock; ++b) {
            const CFResult& cf = r[b];
            norms[b] = cf.norm_phi0;
            for (int i = 0; i < n_iter; ++i) {
                alphas[size_t(b) * n_iter + i] = i < int(cf.alphas.size()) ? cf.alphas[i] : 0.0;
                betas[size_t(b) * n_iter + i] = i < int(cf.betas.size()) ? cf.betas[i] : 0.0;
            }
            if (counts) counts[b] = int(cf.alphas.size());
        }
    });
}

float qkrylov_evaluate_spectral_function(const float* alphas, const float* betas, size_t n,
                                         float norm_phi0, float omega, float E0, float eta) {
    if (!alphas || !betas || n == 0) return 0.0f;
    std::vector<double> a(n), b(n > 0 ? n - 1 : 0);
    for (size_t i = 0; i < n; ++i) a[i] = alphas[i];
    for (size_t i = 0; i + 1 < n; ++i) b[i] = betas[i];
    return float(spectral_function(a.data(), b.data(), n, norm_phi0, omega, E0, eta));
}
double qkrylov_evaluate_spectral_function_f64(const double* alphas, const double* betas, size_t n,
                                              double norm_phi0, double omega, double E0, double eta) {
    if (!alphas || !betas || n == 0) return 0.0;
    return spectral_function(alphas, betas, n, norm_phi0, omega, E0, eta);
}

int qkrylov_hermitian_eigh_f64(int n, const double* a, double* evals, double* vecs) {
    if (n <= 0 || !a || !evals) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        std::vector<cplx> A(size_t(n) * n);
        std::memcpy(A.data(), a, sizeof(cplx) * A.size());
        std::vector<double> ev;
        std::vector<cplx> S;
        hermitian_eigh(A, n, ev, S);
        std::memcpy(evals, ev.data(), sizeof(double) * n);
        if (vecs) std::memcpy(vecs, S.data(), sizeof(cplx) * S.size());
    });
}

int qkrylov_ftlm(qkrylov_hamiltonian_h h, float beta, int n_random, int n_steps,
                 qkrylov_ftlm_result_c_t* result) {
    if (!hok(h) || !result) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const auto r = ftlm_sweep({h->ptr.get()}, {double(beta)}, n_random, n_steps, 0);
        result->beta = float(r[0].beta);
        result->partition_function = float(r[0].Z);
        result->internal_energy = float(r[0].E);
        result->specific_heat = float(r[0].Cv);
    });
}
int qkrylov_ftlm_f64(qkrylov_hamiltonian_h h, double beta, int n_random, int n_steps,
                     qkrylov_ftlm_result_f64_t* result) {
    if (!hok(h) || !result) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        const auto r = ftlm_sweep({h->ptr.get()}, {beta}, n_random, n_steps, 0);
        result->beta = r[0].beta;
        result->partition_function = r[0].Z;
        result->internal_energy = r[0].E;
        result->specific_heat = r[0].Cv;
    });
}
int qkrylov_ftlm_sweep_f64(const qkrylov_hamiltonian_h* hs, int nsec, const double* betas, int nbeta,
                           int n_random, int n_steps, int mode, double* Z, double* E, double* Cv) {
    if (!hs || nsec <= 0 || !betas || nbeta <= 0 || !Z || !E || !Cv || mode < 0 || mode > 3)
        return QKRYLOV_ERROR_INVALID_ARG;
    for (int s = 0; s < nsec; ++s)
        if (!hok(hs[s])) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        std::vector<const Hamiltonian*> v;
        for (int s = 0; s < nsec; ++s) v.push_back(hs[s]->ptr.get());
        const auto r = ftlm_sweep(v, std::vector<double>(betas, betas + nbeta), n_random, n_steps, mode);
        for (size_t i = 0; i < r.size(); ++i) {
            Z[i] = r[i].Z;
            E[i] = r[i].E;
            Cv[i] = r[i].Cv;
        }
    });
}

int qkrylov_reference_start_f64(uint64_t n, unsigned seed, double* out) {
    if (!out && n) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        uniform_start(n, seed, reinterpret_cast<cplx*>(out));
    });
}
int qkrylov_reference_normals_f64(uint64_t n, unsigned seed, double* out) {
    if (!out && n) return QKRYLOV_ERROR_INVALID_ARG;
    return guarded([&] {
        NormalStream s(seed);
        s.fill(reinterpret_cast<cplx*>(out), n);
    });
}

}  // extern "C"
