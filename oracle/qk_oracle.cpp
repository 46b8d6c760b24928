// ORACLE — test infrastructure only.  A CPU restatement of the reference
// qkrylov algorithms, used as the checker for the HIP library.  Only
// tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it;
// nothing in the product links, imports or calls it.
//
// Each routine follows the reference file:line named in its comment.  The
// reference C++ needs Kokkos (fetched from the network by its CMake), so
// this file restates the Kokkos-dependent parts; the Kokkos-free parts
// (bases, sites) are additionally cross-checked against the reference's own
// sources compiled into oracle/_ref (see oracle/Makefile, ref_shim.cpp).
//
// Deviations that cannot change results:
//  * basis lookup uses binary search over the ascending state list instead
//    of std::unordered_map (same index for every state, absent -> absent);
//  * the Lanczos convergence check computes the tridiagonal eigenvalues
//    without accumulating eigenvectors (the QL sweep on d/e does not read z);
//  * ftlm_multi evaluates several temperatures from one Lanczos run per
//    random vector (the Lanczos part of ftlm() does not depend on beta).
// Parity pins: see tests/test_oracle_pins.py.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

using Complex = std::complex<double>;
using StateID = uint64_t;

static thread_local std::string g_err;

// ---------------------------------------------------------------- sector
struct Sector {  // include/qkrylov/symmetry/sector.hpp:10-41
    bool use_sz = false;
    int sz2 = 0;
    bool use_nup = false, use_ndn = false;
    int nup = 0, ndn = 0;
    bool use_n = false;
    int n = 0;
    bool use_nb = false;
    int nb = 0;
};

// ---------------------------------------------------------------- bases
enum Kind { SPIN = 0, FERMION = 1, HUBBARD = 2, TJ = 3 };

struct Basis {
    int kind, N;
    Sector sec;
    std::vector<StateID> states;  // ascending
    size_t size() const { return states.size(); }
    bool contains(StateID s) const { return std::binary_search(states.begin(), states.end(), s); }
    size_t index(StateID s) const {
        auto it = std::lower_bound(states.begin(), states.end(), s);
        if (it == states.end() || *it != s) throw std::runtime_error("State not present in basis");
        return size_t(it - states.begin());
    }
};

static int popcount(uint64_t x) { return __builtin_popcountll(x); }

// src/basis/spinhalf_basis.cpp:12-104, fermion_basis.cpp:12-89,
// hubbard_basis.cpp:12-100, tj_basis.cpp:12-96
static Basis* make_basis(int kind, int N, const Sector& sec) {
    auto* b = new Basis{kind, N, sec, {}};
    if (kind == SPIN || kind == FERMION) {
        if (N <= 0 || N > 63) {
            delete b;
            throw std::runtime_error("N must satisfy 1 <= N <= 63");
        }
        const StateID dim = StateID(1) << N;
        const bool use = (kind == SPIN) ? sec.use_sz : sec.use_n;
        for (StateID s = 0; s < dim; ++s) {
            if (use) {
                if (kind == SPIN) {
                    const int nup = popcount(s);
                    if (nup - (N - nup) != sec.sz2) continue;
                } else if (popcount(s) != sec.n) {
                    continue;
                }
            }
            b->states.push_back(s);
        }
    } else {
        if (N <= 0 || N > 31) {
            delete b;
            throw std::runtime_error("N must satisfy 1 <= N <= 31");
        }
        const StateID dim = StateID(1) << (2 * N);
        StateID up_mask = 0, dn_mask = 0;
        for (int i = 0; i < N; ++i) {
            up_mask |= 1ULL << (2 * i);
            dn_mask |= 1ULL << (2 * i + 1);
        }
        const bool filter = (kind == TJ) || sec.use_nup || sec.use_ndn;
        for (StateID s = 0; s < dim; ++s) {
            if (kind == TJ) {
                bool dbl = false;
                for (int i = 0; i < N; ++i)
                    if (((s >> (2 * i)) & 1ULL) && ((s >> (2 * i + 1)) & 1ULL)) {
                        dbl = true;
                        break;
                    }
                if (dbl) continue;
            }
            if (filter) {
                const bool mu = !sec.use_nup || popcount(s & up_mask) == sec.nup;
                const bool md = !sec.use_ndn || popcount(s & dn_mask) == sec.ndn;
                if (!(mu && md)) continue;
            }
            b->states.push_back(s);
        }
    }
    return b;
}

// ---------------------------------------------------------------- sites
struct LocalAction {  // include/qkrylov/operators/local_action.hpp:9-16
    bool valid = false;
    StateID new_state = 0;
    Complex me = 0.0;
};

static double jw(StateID s, int bitpos) {
    const StateID mask = (1ULL << bitpos) - 1;
    return (popcount(s & mask) % 2 == 0) ? 1.0 : -1.0;
}

// Mutation hook for tests that show the pins catch a planted mistake:
// 1 counts the site's own bit in the Jordan-Wigner string (off by one),
// 2 drops the conj of stored amplitudes.
static int g_mutation = 0;

static double jw_m(StateID s, int bitpos) {
    return g_mutation == 1 ? jw(s, bitpos + 1) : jw(s, bitpos);
}

static LocalAction site_apply(int kind, const std::string& op, int site, StateID s) {
    LocalAction a;
    if (kind == SPIN) {  // src/sites/spinhalf_site.cpp:18-121
        const bool up = (s >> site) & 1ULL;
        if (op == "Sz") { a.valid = true; a.new_state = s; a.me = up ? 0.5 : -0.5; return a; }
        if (op == "Sp") { if (up) return a; a.valid = true; a.new_state = s | (1ULL << site); a.me = 1.0; return a; }
        if (op == "Sm") { if (!up) return a; a.valid = true; a.new_state = s & ~(1ULL << site); a.me = 1.0; return a; }
        if (op == "Sx") { a.valid = true; a.new_state = s ^ (1ULL << site); a.me = 0.5; return a; }
        if (op == "Sy") { a.valid = true; a.new_state = s ^ (1ULL << site); a.me = up ? Complex(0.0, -0.5) : Complex(0.0, 0.5); return a; }
        throw std::runtime_error("Unknown spin operator: " + op);
    }
    if (kind == FERMION) {  // src/sites/fermion_site.cpp:19-84
        const bool occ = (s >> site) & 1ULL;
        const double ph = jw_m(s, site);
        if (op == "N") { a.valid = true; a.new_state = s; a.me = occ ? 1.0 : 0.0; return a; }
        if (op == "Id") { a.valid = true; a.new_state = s; a.me = 1.0; return a; }
        if (op == "C") { if (!occ) return a; a.valid = true; a.new_state = s & ~(1ULL << site); a.me = ph; return a; }
        if (op == "Cdag") { if (occ) return a; a.valid = true; a.new_state = s | (1ULL << site); a.me = ph; return a; }
        throw std::runtime_error("Unknown fermion operator: " + op);
    }
    // src/sites/hubbard_site.cpp:51-125 and tj_site.cpp:47-113
    const bool tj = (kind == TJ);
    const int bu = 2 * site, bd = 2 * site + 1;
    const bool up = (s >> bu) & 1ULL, dn = (s >> bd) & 1ULL;
    const double pu = jw_m(s, bu);
    const double pd = jw_m(s, bd);
    if (op == "Nup") { a.valid = true; a.new_state = s; a.me = up ? 1.0 : 0.0; return a; }
    if (op == "Ndn") { a.valid = true; a.new_state = s; a.me = dn ? 1.0 : 0.0; return a; }
    if (!tj && op == "Nupdn") { a.valid = true; a.new_state = s; a.me = (up && dn) ? 1.0 : 0.0; return a; }
    if (op == "CUp") { if (!up) return a; a.valid = true; a.new_state = s & ~(1ULL << bu); a.me = pu; return a; }
    if (op == "CdagUp") { if (up || (tj && dn)) return a; a.valid = true; a.new_state = s | (1ULL << bu); a.me = pu; return a; }
    if (op == "CDn") { if (!dn) return a; a.valid = true; a.new_state = s & ~(1ULL << bd); a.me = pd; return a; }
    if (op == "CdagDn") { if (dn || (tj && up)) return a; a.valid = true; a.new_state = s | (1ULL << bd); a.me = pd; return a; }
    throw std::runtime_error(std::string(tj ? "Unknown TJ operator: " : "Unknown Hubbard operator: ") + op);
}

// ---------------------------------------------------------------- opsum
struct Factor {
    std::string op;
    int site;
};
struct Term {
    Complex coeff;
    std::vector<Factor> factors;
};

// ---------------------------------------------------------------- CSR H
// src/hamiltonian/matrix_free_hamiltonian.cpp:12-135 (compile) and
// :141-168 (gather apply).  Rows are built in parallel and concatenated in
// row order, which yields the reference's serial CSR exactly.
struct Ham {
    size_t dim = 0;
    std::vector<int64_t> rows;
    std::vector<int64_t> cols;
    std::vector<Complex> vals;
    std::vector<Complex> diag;
};

static void row_entries(const Basis& B, int site_kind, const std::vector<Term>& terms, size_t alpha,
                        std::vector<int64_t>& cols, std::vector<Complex>& vals, Complex& dg) {
    const StateID init = B.states[alpha];
    for (const auto& term : terms) {
        StateID st = init;
        Complex amp = term.coeff;
        bool ok = true;
        for (const auto& f : term.factors) {
            const LocalAction act = site_apply(site_kind, f.op, f.site, st);
            if (!act.valid) {
                ok = false;
                break;
            }
            st = act.new_state;
            amp *= act.me;
        }
        if (!ok || !B.contains(st)) continue;
        const size_t beta = B.index(st);
        const Complex val = g_mutation == 2 ? amp : std::conj(amp);
        cols.push_back(int64_t(beta));
        vals.push_back(val);
        if (beta == alpha) dg += val;
    }
}

static Ham* make_ham(const Basis& B, int site_kind, const std::vector<Term>& terms, size_t row0,
                     size_t row1) {
    auto* H = new Ham();
    H->dim = B.size();
    row1 = std::min(row1, B.size());
    const size_t nr = row1 > row0 ? row1 - row0 : 0;
    H->rows.assign(nr + 1, 0);
    H->diag.assign(nr, Complex(0.0, 0.0));
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    std::vector<std::vector<int64_t>> tc(nt);
    std::vector<std::vector<Complex>> tv(nt);
    std::vector<size_t> tstart(nt + 1, 0);
    for (int t = 0; t <= nt; ++t) tstart[t] = row0 + nr * size_t(t) / size_t(nt);
    std::string err;
#pragma omp parallel num_threads(nt)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        try {
            for (size_t a = tstart[t]; a < tstart[t + 1]; ++a) {
                H->rows[a - row0] = int64_t(tc[t].size());   // local offset for now
                Complex dg(0.0, 0.0);
                row_entries(B, site_kind, terms, a, tc[t], tv[t], dg);
                H->diag[a - row0] = dg;
            }
        } catch (const std::exception& e) {
#pragma omp critical
            err = e.what();
        }
    }
    if (!err.empty()) {
        delete H;
        throw std::runtime_error(err);
    }
    size_t off = 0;
    for (int t = 0; t < nt; ++t) {
        for (size_t a = tstart[t]; a < tstart[t + 1]; ++a) H->rows[a - row0] += int64_t(off);
        off += tc[t].size();
    }
    H->rows[nr] = int64_t(off);
    H->cols.reserve(off);
    H->vals.reserve(off);
    for (int t = 0; t < nt; ++t) {
        H->cols.insert(H->cols.end(), tc[t].begin(), tc[t].end());
        H->vals.insert(H->vals.end(), tv[t].begin(), tv[t].end());
    }
    return H;
}

static void ham_apply(const Ham& H, const Complex* x, Complex* y) {
    const int64_t nr = int64_t(H.rows.size()) - 1;
#pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < nr; ++a) {
        Complex sum(0.0, 0.0);
        for (int64_t j = H.rows[a]; j < H.rows[a + 1]; ++j) sum += H.vals[j] * x[H.cols[j]];
        y[a] = sum;
    }
}

// ---------------------------------------------------------------- BLAS-1
// include/qkrylov/linalg/vector_ops.hpp:14-137 (sequential sums)
static Complex dot(const std::vector<Complex>& x, const std::vector<Complex>& y) {
    Complex s(0.0, 0.0);
    for (size_t i = 0; i < x.size(); ++i) s += std::conj(x[i]) * y[i];
    return s;
}
static double norm(const std::vector<Complex>& x) { return std::sqrt(dot(x, x).real()); }
static void axpy(Complex a, const std::vector<Complex>& x, std::vector<Complex>& y) {
    for (size_t i = 0; i < x.size(); ++i) y[i] += a * x[i];
}
static void scal(Complex a, std::vector<Complex>& x) {
    for (auto& v : x) v *= a;
}
static void normalize(std::vector<Complex>& x) {
    const double n = norm(x);
    if (n == 0.0) return;
    scal(Complex(1.0 / n, 0.0), x);
}

// ---------------------------------------------------------------- QL
// src/solvers/lanczos.cpp:25-98 and ftlm.cpp:61-119 (identical sweeps)
static void ql(const std::vector<double>& alpha, const std::vector<double>& beta, int n,
               std::vector<double>& dout, std::vector<std::vector<double>>* zout) {
    std::vector<double> d(alpha.begin(), alpha.begin() + n);
    std::vector<double> e(beta);
    e.resize(std::max<size_t>(e.size(), size_t(n)), 0.0);
    std::vector<std::vector<double>> z;
    if (zout) {
        z.assign(n, std::vector<double>(n, 0.0));
        for (int i = 0; i < n; ++i) z[i][i] = 1.0;
    }
    for (int iter = 0; iter < 1000; ++iter) {
        for (int i = 0; i < n - 1; ++i)
            if (std::abs(e[i]) < 1e-14 * (std::abs(d[i]) + std::abs(d[i + 1]))) e[i] = 0.0;
        int m = n - 1;
        while (m > 0 && e[m - 1] == 0.0) m--;
        if (m == 0) break;
        int l = m - 1;
        while (l > 0 && e[l - 1] != 0.0) l--;
        double b = (d[m - 1] - d[m]) / 2.0;
        double c = e[m - 1] * e[m - 1];
        double s = std::sqrt(b * b + c);
        double shift = (b > 0) ? d[m] - c / (b + s) : d[m] - c / (b - s);
        double p = d[l] - shift;
        double g = e[l];
        for (int i = l; i < m; ++i) {
            double r = std::hypot(p, g);
            double cs = p / r, sn = g / r;
            if (i > l) e[i - 1] = r;
            double f = cs * d[i] + sn * e[i];
            double gn = cs * e[i] + sn * d[i + 1];
            double h = sn * d[i] - cs * e[i];
            double k = sn * e[i] - cs * d[i + 1];
            d[i] = cs * f + sn * gn;
            e[i] = cs * h + sn * k;
            d[i + 1] = sn * h - cs * k;
            if (zout)
                for (int j = 0; j < n; ++j) {
                    double z1 = z[j][i], z2 = z[j][i + 1];
                    z[j][i] = cs * z1 + sn * z2;
                    z[j][i + 1] = sn * z1 - cs * z2;
                }
            if (i < m - 1) {
                p = e[i];
                g = sn * e[i + 1];
                e[i + 1] = -cs * e[i + 1];
            }
        }
    }
    dout = d;
    if (zout) *zout = z;
}

static double ground(const std::vector<double>& a, const std::vector<double>& b, int n,
                     std::vector<double>* vec) {
    if (n == 0) return 0.0;
    if (n == 1) {
        if (vec) *vec = {1.0};
        return a[0];
    }
    std::vector<double> d;
    std::vector<std::vector<double>> z;
    ql(a, b, n, d, vec ? &z : nullptr);
    int mi = 0;
    for (int i = 1; i < n; ++i)
        if (d[i] < d[mi]) mi = i;
    if (vec) {
        vec->resize(n);
        for (int i = 0; i < n; ++i) (*vec)[i] = z[i][mi];
    }
    return d[mi];
}

// ---------------------------------------------------------------- Lanczos
// src/solvers/lanczos.cpp:101-202.  start: null for the reference start
// vector (mt19937(1234), uniform(-1,1), imaginary part drawn first).
struct LanczosOut {
    double energy = 0.0;
    int iterations = 0;
    bool converged = false;
    std::vector<Complex> vec;
};

static std::vector<Complex> uniform_start(size_t n, unsigned seed) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> dist(-1.0, 1.0);
    std::vector<Complex> v(n);
    for (size_t i = 0; i < n; ++i) {
        const double im = dist(rng);
        const double re = dist(rng);
        v[i] = Complex(re, im);
    }
    return v;
}

static LanczosOut lanczos(const Ham& H, int maxiter, double tol, const Complex* start) {
    LanczosOut res;
    const size_t dim = H.dim;
    if (dim == 0) return res;
    std::vector<Complex> v_curr = start ? std::vector<Complex>(start, start + dim) : uniform_start(dim, 1234);
    std::vector<Complex> v_prev(dim), w(dim);
    normalize(v_curr);
    std::vector<std::vector<Complex>> basis{v_curr};
    std::vector<double> alphas, betas;
    double e_old = 1e100;
    const int nmax = int(std::min<int64_t>(maxiter, int64_t(dim)));
    for (int iter = 0; iter < nmax; ++iter) {
        res.iterations = iter + 1;
        ham_apply(H, v_curr.data(), w.data());
        const double alpha = dot(v_curr, w).real();
        alphas.push_back(alpha);
        axpy(-alpha, v_curr, w);
        if (iter > 0) axpy(-betas.back(), v_prev, w);
        for (const auto& bv : basis) axpy(-dot(bv, w), bv, w);
        const double beta = norm(w);
        if (beta < 1e-15) {
            res.converged = true;
            break;
        }
        if (iter + 1 == nmax) break;
        betas.push_back(beta);
        v_prev = v_curr;
        v_curr = w;
        scal(1.0 / beta, v_curr);
        basis.push_back(v_curr);
        if (iter > 0) {
            const double e = ground(alphas, betas, int(alphas.size()), nullptr);
            if (std::abs(e - e_old) < tol) {
                res.converged = true;
                break;
            }
            e_old = e;
        }
    }
    std::vector<double> z;
    res.energy = ground(alphas, betas, int(alphas.size()), &z);
    res.vec.assign(dim, Complex(0.0, 0.0));
    for (size_t i = 0; i < alphas.size(); ++i) axpy(Complex(z[i], 0.0), basis[i], res.vec);
    normalize(res.vec);
    return res;
}

// ---------------------------------------------------------------- Davidson
// src/solvers/davidson.cpp:22-82 (Jacobi, restated as written, including
// its update A[p][i] = A[i][p], which is exact only for real matrices) and
// :85-227 (driver).
static void jacobi(const std::vector<std::vector<Complex>>& Hs, std::vector<double>& evals,
                   std::vector<std::vector<Complex>>& evecs) {
    const int n = int(Hs.size());
    auto A = Hs;
    std::vector<std::vector<Complex>> V(n, std::vector<Complex>(n, 0.0));
    for (int i = 0; i < n; ++i) V[i][i] = 1.0;
    for (int iter = 0; iter < 1000; ++iter) {
        double max_off = 0.0;
        int p = 0, q = 0;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j)
                if (std::abs(A[i][j]) > max_off) {
                    max_off = std::abs(A[i][j]);
                    p = i;
                    q = j;
                }
        if (max_off < 1e-15) break;
        Complex app = A[p][p], aqq = A[q][q], apq = A[p][q];
        double phi = 0.5 * std::atan2(2.0 * std::abs(apq), std::real(app - aqq));
        Complex c = std::cos(phi);
        Complex s = std::sin(phi) * apq / std::abs(apq);
        for (int i = 0; i < n; ++i) {
            Complex vip = V[i][p], viq = V[i][q];
            V[i][p] = c * vip + s * viq;
            V[i][q] = -std::conj(s) * vip + c * viq;
            if (i != p && i != q) {
                Complex aip = A[i][p], aiq = A[i][q];
                A[i][p] = A[p][i] = c * aip + s * aiq;
                A[i][q] = A[q][i] = -std::conj(s) * aip + c * aiq;
            }
        }
        A[p][p] = c * c * app + s * std::conj(s) * aqq + 2.0 * std::real(c * s * std::conj(apq));
        A[q][q] = s * std::conj(s) * app + c * c * aqq - 2.0 * std::real(c * s * std::conj(apq));
        A[p][q] = A[q][p] = 0.0;
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int i, int j) { return std::real(A[i][i]) < std::real(A[j][j]); });
    evals.clear();
    evecs.clear();
    for (int i : idx) {
        evals.push_back(std::real(A[i][i]));
        std::vector<Complex> v(n);
        for (int j = 0; j < n; ++j) v[j] = V[j][i];
        evecs.push_back(v);
    }
}

struct DavidsonOut {
    std::vector<double> evals;
    std::vector<std::vector<Complex>> evecs;
    int iterations = 0;
    bool converged = false;
};

static DavidsonOut davidson(const Ham& H, int n_eig, int max_subspace, double tol) {
    DavidsonOut res;
    const size_t dim = H.dim;
    if (dim == 0) return res;
    n_eig = int(std::min<int64_t>(n_eig, int64_t(dim)));
    max_subspace = std::max(max_subspace, 2 * n_eig);
    std::vector<size_t> idx(dim);
    std::iota(idx.begin(), idx.end(), size_t(0));
    std::sort(idx.begin(), idx.end(), [&](size_t i, size_t j) { return H.diag[i].real() < H.diag[j].real(); });
    std::mt19937 rng(1234);
    std::uniform_real_distribution<double> dist(-1.0, 1.0);
    std::vector<std::vector<Complex>> V, HV;
    for (int i = 0; i < n_eig; ++i) {
        std::vector<Complex> v(dim, 0.0);
        v[idx[i]] = 1.0;
        for (const auto& vp : V) axpy(-dot(vp, v), vp, v);
        double nrm = norm(v);
        if (nrm < 1e-10) {
            for (size_t j = 0; j < dim; ++j) {
                const double im = dist(rng);
                const double re = dist(rng);
                v[j] = Complex(re, im);
            }
            for (const auto& vp : V) axpy(-dot(vp, v), vp, v);
            nrm = norm(v);
        }
        scal(1.0 / nrm, v);
        V.push_back(v);
        std::vector<Complex> hv(dim);
        ham_apply(H, v.data(), hv.data());
        HV.push_back(hv);
    }
    for (int iter = 0; iter < 100; ++iter) {
        const int m = int(V.size());
        std::vector<std::vector<Complex>> Hs(m, std::vector<Complex>(m));
        for (int i = 0; i < m; ++i)
            for (int j = 0; j <= i; ++j) {
                Hs[i][j] = dot(V[i], HV[j]);
                if (i != j) Hs[j][i] = std::conj(Hs[i][j]);
            }
        std::vector<double> ev;
        std::vector<std::vector<Complex>> es;
        jacobi(Hs, ev, es);
        bool all = true;
        std::vector<std::vector<Complex>> corr;
        for (int k = 0; k < n_eig; ++k) {
            const double lam = ev[k];
            const auto& s = es[k];
            std::vector<Complex> r(dim, 0.0);
            for (int i = 0; i < m; ++i) {
                axpy(s[i], HV[i], r);
                axpy(Complex(-lam * s[i].real(), -lam * s[i].imag()), V[i], r);
            }
            if (norm(r) > tol) {
                all = false;
                std::vector<Complex> t(dim);
                for (size_t i = 0; i < dim; ++i) {
                    Complex diff = H.diag[i] - lam;
                    if (std::abs(diff) < 1e-10) diff = (diff.real() >= 0) ? 1e-10 : -1e-10;
                    t[i] = r[i] / diff;
                }
                corr.push_back(t);
            }
        }
        if (all || iter == 99) {
            res.iterations = iter + 1;
            res.converged = all;
            for (int k = 0; k < n_eig; ++k) {
                res.evals.push_back(ev[k]);
                std::vector<Complex> e(dim, 0.0);
                for (int i = 0; i < m; ++i) axpy(es[k][i], V[i], e);
                res.evecs.push_back(e);
            }
            return res;
        }
        if (V.size() + corr.size() > size_t(max_subspace)) {
            std::vector<std::vector<Complex>> nV, nHV;
            for (int k = 0; k < n_eig; ++k) {
                std::vector<Complex> rv(dim, 0.0), rh(dim, 0.0);
                for (int i = 0; i < m; ++i) {
                    axpy(es[k][i], V[i], rv);
                    axpy(es[k][i], HV[i], rh);
                }
                nV.push_back(rv);
                nHV.push_back(rh);
            }
            V = nV;
            HV = nHV;
        }
        for (auto& t : corr) {
            for (const auto& v : V) axpy(-dot(v, t), v, t);
            const double nrm = norm(t);
            if (nrm > 1e-10) {
                scal(1.0 / nrm, t);
                V.push_back(t);
                std::vector<Complex> ht(dim);
                ham_apply(H, t.data(), ht.data());
                HV.push_back(ht);
            }
        }
    }
    return DavidsonOut{};
}

// ---------------------------------------------------------------- tridiag
// src/solvers/ftlm.cpp:24-53 and dynamics.cpp:12-59 (same recurrence)
static void recurrence(const Ham& H, std::vector<Complex> v_curr, int n_steps,
                       std::vector<double>& alphas, std::vector<double>& betas) {
    const size_t dim = H.dim;
    std::vector<Complex> v_prev(dim), w(dim);
    for (int i = 0; i < n_steps; ++i) {
        ham_apply(H, v_curr.data(), w.data());
        const double alpha = dot(v_curr, w).real();
        alphas.push_back(alpha);
        axpy(-alpha, v_curr, w);
        if (i > 0) axpy(-betas.back(), v_prev, w);
        const double beta = norm(w);
        if (beta < 1e-15) break;
        betas.push_back(beta);
        v_prev = v_curr;
        v_curr = w;
        scal(1.0 / beta, v_curr);
    }
}

// ftlm() for several betas at once (ftlm.cpp:122-170)
static void ftlm_multi(const Ham& H, const double* betas, int nb, int n_random, int n_steps,
                       double* Z, double* E, double* Cv) {
    const size_t dim = H.dim;
    std::vector<double> zs(nb, 0.0), es(nb, 0.0), e2s(nb, 0.0);
    if (dim == 0) {
        for (int b = 0; b < nb; ++b) Z[b] = E[b] = Cv[b] = 0.0;
        return;
    }
    std::mt19937 rng(42);
    std::normal_distribution<double> dist(0.0, 1.0);
    for (int r = 0; r < n_random; ++r) {
        std::vector<Complex> rv(dim);
        for (size_t i = 0; i < dim; ++i) {
            const double im = dist(rng);
            const double re = dist(rng);
            rv[i] = Complex(re, im);
        }
        const double nrm = norm(rv);
        std::vector<Complex> v = rv;
        normalize(v);
        std::vector<double> al, be;
        recurrence(H, v, n_steps, al, be);
        const int n = int(al.size());
        std::vector<double> ev;
        std::vector<double> z0;
        if (n == 1) {
            ev = {al[0]};
            z0 = {1.0};
        } else if (n > 1) {
            std::vector<std::vector<double>> z;
            ql(al, be, n, ev, &z);
            z0.resize(n);
            for (int i = 0; i < n; ++i) z0[i] = z[0][i];
        }
        for (int b = 0; b < nb; ++b)
            for (int i = 0; i < n; ++i) {
                const double w = nrm * nrm * z0[i] * z0[i] * std::exp(-betas[b] * ev[i]);
                zs[b] += w;
                es[b] += ev[i] * w;
                e2s[b] += ev[i] * ev[i] * w;
            }
    }
    for (int b = 0; b < nb; ++b) {
        const double z = zs[b] / n_random, e = es[b] / n_random, e2 = e2s[b] / n_random;
        Z[b] = z;
        E[b] = e / z;
        Cv[b] = betas[b] * betas[b] * (e2 / z - (e / z) * (e / z));
    }
}

// dynamics.cpp:61-89
static double spectral(const double* alphas, const double* betas, size_t n, double norm_phi0,
                       double omega, double E0, double eta) {
    if (n == 0) return 0.0;
    std::complex<double> z(omega + E0, eta);
    std::complex<double> f = 0.0;
    for (int i = int(n) - 1; i >= 0; --i) {
        if (i == int(n) - 1) f = 1.0 / (z - alphas[i]);
        else f = 1.0 / (z - alphas[i] - betas[i] * betas[i] * f);
    }
    return -1.0 / M_PI * std::imag(norm_phi0 * norm_phi0 * f);
}

}  // namespace orc

// ======================================================================
// C ABI for the tests (ctypes)
// ======================================================================
using namespace orc;

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }
void orc_set_mutation(int m) { g_mutation = m; }
int orc_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void* orc_basis_create(int kind, int N, int use_sz, int sz2, int use_nup, int nup, int use_ndn,
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
        return make_basis(kind, N, s);
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
// A basis from an explicit ascending list of states: for states wider than
// 32 bits, where the reference's loop over all 2^nbits states (make_basis)
// is out of reach, the caller lists the sector's states itself.  Only the
// state list and the site kind matter to the Hamiltonian built on it.
void* orc_basis_from_states(int kind, int N, const uint64_t* states, uint64_t n) {
    try {
        auto* b = new Basis{kind, N, Sector{}, {}};
        b->states.assign(states, states + n);
        for (uint64_t i = 1; i < n; ++i)
            if (b->states[i] <= b->states[i - 1]) {
                delete b;
                throw std::runtime_error("states must be strictly ascending");
            }
        return b;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
void orc_basis_destroy(void* b) { delete static_cast<Basis*>(b); }
uint64_t orc_basis_dim(void* b) { return static_cast<Basis*>(b)->size(); }
void orc_basis_states(void* b, uint64_t* out) {
    auto* B = static_cast<Basis*>(b);
    std::memcpy(out, B->states.data(), sizeof(uint64_t) * B->size());
}
int64_t orc_basis_index(void* b, uint64_t s) {
    auto* B = static_cast<Basis*>(b);
    return B->contains(s) ? int64_t(B->index(s)) : -1;
}

int orc_site_apply(int kind, const char* op, int site, uint64_t s, uint64_t* new_state,
                   double* re, double* im) {
    try {
        const LocalAction a = site_apply(kind, op, site, s);
        *new_state = a.new_state;
        *re = a.me.real();
        *im = a.me.imag();
        return a.valid ? 1 : 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

static std::vector<Term> read_terms(int nterms, const double* cre, const double* cim, const int* nfac,
                                    const char** ops, const int* sites) {
    std::vector<Term> t(nterms);
    int k = 0;
    for (int i = 0; i < nterms; ++i) {
        t[i].coeff = Complex(cre[i], cim[i]);
        for (int f = 0; f < nfac[i]; ++f, ++k) t[i].factors.push_back({ops[k], sites[k]});
    }
    return t;
}

// rows [row0, row1) of H (row1 = -1: all)
void* orc_ham_create(void* basis, int site_kind, int nterms, const double* cre, const double* cim,
                     const int* nfac, const char** ops, const int* sites, int64_t row0, int64_t row1) {
    try {
        const Basis& B = *static_cast<Basis*>(basis);
        const auto terms = read_terms(nterms, cre, cim, nfac, ops, sites);
        return make_ham(B, site_kind, terms, size_t(row0), row1 < 0 ? B.size() : size_t(row1));
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}
void orc_ham_destroy(void* h) { delete static_cast<Ham*>(h); }
uint64_t orc_ham_nnz(void* h) { return static_cast<Ham*>(h)->cols.size(); }
uint64_t orc_ham_nrows(void* h) { return static_cast<Ham*>(h)->rows.size() - 1; }
void orc_ham_csr(void* h, int64_t* rows, int64_t* cols, double* vals) {
    auto* H = static_cast<Ham*>(h);
    std::memcpy(rows, H->rows.data(), sizeof(int64_t) * H->rows.size());
    std::memcpy(cols, H->cols.data(), sizeof(int64_t) * H->cols.size());
    std::memcpy(vals, H->vals.data(), sizeof(Complex) * H->vals.size());
}
void orc_ham_diag(void* h, double* out) {
    auto* H = static_cast<Ham*>(h);
    std::memcpy(out, H->diag.data(), sizeof(Complex) * H->diag.size());
}
// y[rows of h] = H x (x of full dimension)
void orc_ham_apply(void* h, const double* x, double* y) {
    ham_apply(*static_cast<Ham*>(h), reinterpret_cast<const Complex*>(x), reinterpret_cast<Complex*>(y));
}

int orc_lanczos(void* h, int maxiter, double tol, const double* start, double* energy, int* iters,
                int* conv, double* vec) {
    try {
        const auto r = lanczos(*static_cast<Ham*>(h), maxiter, tol, reinterpret_cast<const Complex*>(start));
        *energy = r.energy;
        *iters = r.iterations;
        *conv = r.converged;
        if (vec) std::memcpy(vec, r.vec.data(), sizeof(Complex) * r.vec.size());
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -2;
    }
}

int orc_davidson(void* h, int n_eig, int max_subspace, double tol, double* evals, double* evecs,
                 int* iters, int* conv) {
    try {
        const auto* H = static_cast<Ham*>(h);
        const auto r = davidson(*H, n_eig, max_subspace, tol);
        for (size_t k = 0; k < r.evals.size(); ++k) {
            evals[k] = r.evals[k];
            if (evecs) std::memcpy(evecs + 2 * k * H->dim, r.evecs[k].data(), sizeof(Complex) * H->dim);
        }
        *iters = r.iterations;
        *conv = r.converged;
        return int(r.evals.size());
    } catch (const std::exception& e) {
        g_err = e.what();
        return -2;
    }
}

int orc_cf(void* h, const double* phi0, int n_iter, double* alphas, double* betas, double* norm_out) {
    const auto* H = static_cast<Ham*>(h);
    const size_t dim = H->dim;
    if (dim == 0) {
        *norm_out = 0.0;
        return 0;
    }
    std::vector<Complex> p(reinterpret_cast<const Complex*>(phi0), reinterpret_cast<const Complex*>(phi0) + dim);
    const double np = norm(p);
    *norm_out = np;
    if (np < 1e-15) {
        *norm_out = 0.0;
        return 0;
    }
    scal(1.0 / np, p);
    std::vector<double> al, be;
    recurrence(*H, p, n_iter, al, be);
    for (size_t i = 0; i < al.size(); ++i) alphas[i] = al[i];
    for (size_t i = 0; i < be.size(); ++i) betas[i] = be[i];
    return int(al.size());
}

double orc_spectral(const double* alphas, const double* betas, uint64_t n, double norm_phi0,
                    double omega, double E0, double eta) {
    return spectral(alphas, betas, n, norm_phi0, omega, E0, eta);
}

void orc_ftlm(void* h, const double* betas, int nb, int n_random, int n_steps, double* Z, double* E,
              double* Cv) {
    ftlm_multi(*static_cast<Ham*>(h), betas, nb, n_random, n_steps, Z, E, Cv);
}

void orc_uniform_start(uint64_t n, unsigned seed, double* out) {
    const auto v = uniform_start(n, seed);
    std::memcpy(out, v.data(), sizeof(Complex) * n);
}
void orc_normal_stream(uint64_t n, unsigned seed, double* out) {
    std::mt19937 rng(seed);
    std::normal_distribution<double> dist(0.0, 1.0);
    for (uint64_t i = 0; i < n; ++i) {
        const double im = dist(rng);
        const double re = dist(rng);
        out[2 * i] = re;
        out[2 * i + 1] = im;
    }
}

}  // extern "C"
