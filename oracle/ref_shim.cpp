// ORACLE (reference build) — test infrastructure only.  A C-ABI driver
// around the reference's own Kokkos-free sources (src/basis/*.cpp,
// src/sites/*.cpp, src/operators/opsum.cpp), compiled from where they lie
// under the reference tree by oracle/Makefile into oracle/_ref/.  Used only
// to pin the oracle's bases, site actions and CSR rows against the
// reference's real classes.  The CSR loop below drives the reference's
// Basis/Site/OpSum objects the way MatrixFreeHamiltonian's constructor does
// (src/hamiltonian/matrix_free_hamiltonian.cpp:45-91), which itself needs
// Kokkos and cannot be built here.
#include "qkrylov/basis/fermion_basis.hpp"
#include "qkrylov/basis/hubbard_basis.hpp"
#include "qkrylov/basis/spinhalf_basis.hpp"
#include "qkrylov/basis/tj_basis.hpp"
#include "qkrylov/operators/opsum.hpp"
#include "qkrylov/sites/fermion_site.hpp"
#include "qkrylov/sites/hubbard_site.hpp"
#include "qkrylov/sites/spinhalf_site.hpp"
#include "qkrylov/sites/tj_site.hpp"

#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace qkrylov;
using namespace qkrylov::QKRYLOV_PRECISION_NAMESPACE;

namespace {
thread_local std::string g_err;

std::shared_ptr<Site> make_site(int kind) {
    switch (kind) {
        case 0: return std::make_shared<SpinHalfSite>();
        case 1: return std::make_shared<FermionSite>();
        case 2: return std::make_shared<HubbardSite>();
        default: return std::make_shared<TJSite>();
    }
}

struct RefBasis {
    std::shared_ptr<Basis> b;
};

struct RefCsr {
    std::vector<int64_t> rows, cols;
    std::vector<Complex> vals, diag;
};
}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

void* ref_basis_create(int kind, int N, int use_sz, int sz2, int use_nup, int nup, int use_ndn,
                       int ndn, int use_n, int n) {
    try {
        Sector s;
        s.use_sz = use_sz;
        s.sz2 = sz2;
        s.use_nup = use_nup;
        s.nup = nup;
        s.use_ndn = use_ndn;
        s.ndn = ndn;
        s.use_n = use_n;
        s.n = n;
        auto* r = new RefBasis();
        switch (kind) {
            case 0: r->b = std::make_shared<SpinHalfBasis>(N, s); break;
            case 1: r->b = std::make_shared<FermionBasis>(N, s); break;
            case 2: r->b = std::make_shared<HubbardBasis>(N, s); break;
            default: r->b = std::make_shared<TJBasis>(N, s); break;
        }
        return r;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
void ref_basis_destroy(void* b) { delete static_cast<RefBasis*>(b); }
uint64_t ref_basis_dim(void* b) { return static_cast<RefBasis*>(b)->b->size(); }
void ref_basis_states(void* b, uint64_t* out) {
    auto& B = *static_cast<RefBasis*>(b)->b;
    for (Index i = 0; i < B.size(); ++i) out[i] = B.state(i);
}
int64_t ref_basis_index(void* b, uint64_t s) {
    auto& B = *static_cast<RefBasis*>(b)->b;
    if (!B.contains(s)) return -1;
    return int64_t(B.index(s));
}

int ref_site_apply(int kind, const char* op, int site, uint64_t s, uint64_t* new_state, double* re,
                   double* im) {
    try {
        const auto a = make_site(kind)->apply(op, site, s);
        *new_state = a.new_state;
        *re = a.matrix_element.real();
        *im = a.matrix_element.imag();
        return a.valid ? 1 : 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

void* ref_csr_create(void* basis, int site_kind, int nterms, const double* cre, const double* cim,
                     const int* nfac, const char** ops, const int* sites) {
    try {
        auto& B = *static_cast<RefBasis*>(basis)->b;
        auto site = make_site(site_kind);
        OpSum os;
        int k = 0;
        for (int i = 0; i < nterms; ++i) {
            OperatorTerm t;
            t.coeff = Complex(cre[i], cim[i]);
            for (int f = 0; f < nfac[i]; ++f, ++k) t.factors.push_back({std::string(ops[k]), sites[k]});
            os.add_term(t);
        }
        auto* C = new RefCsr();
        const Index dim = B.size();
        C->rows.assign(dim + 1, 0);
        C->diag.assign(dim, Complex(0.0, 0.0));
        for (Index a = 0; a < dim; ++a) {
            C->rows[a] = int64_t(C->cols.size());
            const StateID s0 = B.state(a);
            for (const auto& term : os.terms()) {
                StateID st = s0;
                Complex amp = term.coeff;
                bool ok = true;
                for (const auto& f : term.factors) {
                    auto act = site->apply(f.op, f.site, st);
                    if (!act.valid) {
                        ok = false;
                        break;
                    }
                    st = act.new_state;
                    amp *= act.matrix_element;
                }
                if (!ok || !B.contains(st)) continue;
                const Index beta = B.index(st);
                const Complex val(amp.real(), -amp.imag());
                C->cols.push_back(int64_t(beta));
                C->vals.push_back(val);
                if (beta == a) C->diag[a] += val;
            }
        }
        C->rows[dim] = int64_t(C->cols.size());
        return C;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
void ref_csr_destroy(void* c) { delete static_cast<RefCsr*>(c); }
uint64_t ref_csr_nnz(void* c) { return static_cast<RefCsr*>(c)->cols.size(); }
void ref_csr_get(void* c, int64_t* rows, int64_t* cols, double* vals, double* diag) {
    auto* C = static_cast<RefCsr*>(c);
    std::memcpy(rows, C->rows.data(), sizeof(int64_t) * C->rows.size());
    std::memcpy(cols, C->cols.data(), sizeof(int64_t) * C->cols.size());
    std::memcpy(vals, C->vals.data(), sizeof(Complex) * C->vals.size());
    std::memcpy(diag, C->diag.data(), sizeof(Complex) * C->diag.size());
}

}  // extern "C"
