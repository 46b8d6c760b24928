// Krylov solver drivers.  Vectors live in HBM; the host only sees the
// Krylov scalars (alpha, beta, projected matrices) and runs the small
// eigensolves.  Reference semantics are cited per routine.
#include "qk_solvers.hpp"

#include "qk_hostla.hpp"
#include "qk_rng.hpp"
#include "qk_vec.hpp"
#include "qk_threads.hpp"
#include "qk_store.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>

namespace qk {

namespace {

// Phase timer, printed to stderr when QKRYLOV_TIMING is set.
struct Timer {
    const char* name;
    bool on;
    std::chrono::steady_clock::time_point t0;
    explicit Timer(const char* n)
        : name(n), on(std::getenv("QKRYLOV_TIMING") != nullptr), t0(std::chrono::steady_clock::now()) {}
    void mark(const char* what) {
        if (!on) return;
        const auto t = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[qkrylov %s] %-24s %9.3f ms\n", name, what,
                     std::chrono::duration<double, std::milli>(t - t0).count());
        t0 = t;
    }
};

// Per-phase totals of a solver loop (QKRYLOV_TIMING): each mark waits for
// the stream and charges the time since the previous mark to `what`.
struct PhaseClock {
    const char* name;
    bool on;
    std::chrono::steady_clock::time_point t0;
    std::vector<std::pair<std::string, double>> acc;
    int extra = 0;   // solver-specific count reported with the totals
    explicit PhaseClock(const char* n)
        : name(n), on(std::getenv("QKRYLOV_TIMING") != nullptr), t0(std::chrono::steady_clock::now()) {}
    void mark(hipStream_t st, const char* what) {
        if (!on) return;
        QK_HIP(hipStreamSynchronize(st));
        const auto t = std::chrono::steady_clock::now();
        const double ms = std::chrono::duration<double, std::milli>(t - t0).count();
        t0 = t;
        for (auto& a : acc)
            if (a.first == what) {
                a.second += ms;
                return;
            }
        acc.emplace_back(what, ms);
    }
    ~PhaseClock() {
        if (!on) return;
        for (auto& a : acc) std::fprintf(stderr, "[qkrylov %s] %-24s %9.3f ms\n", name, a.first.c_str(), a.second);
        std::fprintf(stderr, "[qkrylov %s] count %d\n", name, extra);
    }
};

// The reference normal stream for a seed (std::mt19937 + normal(0,1),
// ftlm.cpp:133-143) in pinned host memory, kept for the process: a longer
// request extends it (the stream for n draws is a prefix of the stream for
// more), so repeated FTLM calls neither regenerate nor restage it.
std::shared_ptr<const cplx> reference_normals(unsigned seed, size_t n) {
    static std::mutex m;
    // never destroyed: the last buffer is released by process exit, not by a
    // static destructor that could run after the HIP runtime is torn down
    static auto* cached = new std::shared_ptr<const cplx>();
    static unsigned cseed = 0;
    static size_t cn = 0;
    std::lock_guard<std::mutex> lk(m);
    if (*cached && cseed == seed && cn >= n) return *cached;
    cplx* p = nullptr;
    QK_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), sizeof(cplx) * std::max<size_t>(n, 1)));
    NormalStream ns(seed);
    ns.fill(p, n);
    *cached = std::shared_ptr<const cplx>(p, [](const cplx* q) { (void)hipHostFree(const_cast<cplx*>(q)); });
    cseed = seed;
    cn = n;
    return *cached;
}

// The same stream in device memory of the current device, for streams of at
// most 1 GiB (null beyond: callers stage from the host copy), kept for the
// process like the host copy, so a repeated sweep copies its start vectors
// device to device instead of over PCIe (the N=20 FTLM sweep: 1.07 GB of
// start vectors from a 189 MB stream).
const cplx* reference_normals_dev(unsigned seed, size_t n, const cplx* host, hipStream_t st) {
    if (sizeof(cplx) * n > (size_t(1) << 30)) return nullptr;
    static std::mutex m;
    static auto* cached = new DevBuf();
    static unsigned cseed = 0;
    static size_t cn = 0;
    static int cdev = -1;
    int dev = 0;
    QK_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(m);
    if (cached->get() && cseed == seed && cn >= n && cdev == dev) return cached->get<cplx>();
    *cached = DevBuf(sizeof(cplx) * std::max<size_t>(n, 1));
    QK_HIP(hipMemcpyAsync(cached->get(), host, sizeof(cplx) * n, hipMemcpyHostToDevice, st));
    QK_HIP(hipStreamSynchronize(st));
    cseed = seed;
    cn = n;
    cdev = dev;
    return cached->get<cplx>();
}

template <typename T>
T fetch(const void* dev, hipStream_t st) {
    T v;
    QK_HIP(hipMemcpyAsync(&v, dev, sizeof(T), hipMemcpyDeviceToHost, st));
    QK_HIP(hipStreamSynchronize(st));
    return v;
}


double dev_norm(const cplx* x, u64 n, DevBuf& red, DevBuf& sc, hipStream_t st) {
    vec::norm2(x, n, sc.get<double>(), red, st);
    return std::sqrt(fetch<double>(sc.get<double>(), st));
}

// f(i) for i < n on the host threads, item i on thread slot i % nt
template <typename F>
void parallel_for(int n, F&& f) {
    if (n <= 0) return;
    const unsigned nt = host_threads(u64(n));
    run_ranges(nt, nt, [&](u64 a, u64 b) {
        for (u64 t = a; t < b; ++t)
            for (int i = int(t); i < n; i += int(nt)) f(i);
    });
}

// Two classical Gram-Schmidt passes of w against k vectors (pointer array
// on the device).  Replaces the reference's single modified Gram-Schmidt
// sweep (lanczos.cpp:149-151); both leave w orthogonal to the same span.
void cgs2(const cplx* const* ptrs, int k, cplx* w, u64 n, DevBuf& h, DevBuf& red, hipStream_t st) {
    if (k <= 0) return;
    h.resize(sizeof(cplx) * k);
    for (int pass = 0; pass < 2; ++pass) {
        vec::mdot(ptrs, k, w, n, h.get<cplx>(), red, st);
        vec::msub(ptrs, k, h.get<cplx>(), w, n, st);
    }
}
void cgs2(const DevBuf& ptrs, int k, cplx* w, u64 n, DevBuf& h, DevBuf& red, hipStream_t st) {
    cgs2(ptrs.get<const cplx*>(), k, w, n, h, red, st);
}

}  // namespace

// ---------------------------------------------------------------------------
// Lanczos ground state (reference src/solvers/lanczos.cpp:101-202)
// ---------------------------------------------------------------------------
namespace {

// A captured one-step graph: replayed once per step (all indices on the device).
struct StepGraph {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    ~StepGraph() {
        if (ge) (void)hipGraphExecDestroy(ge);
        if (g) (void)hipGraphDestroy(g);
    }
    template <typename F>
    void capture(hipStream_t st, F&& body) {
        QK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            body();
        } catch (...) {
            (void)hipStreamEndCapture(st, &g);
            throw;
        }
        QK_HIP(hipStreamEndCapture(st, &g));
        QK_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    }
};

// pinned host memory (step records read back while later steps run)
struct PinnedBuf {
    void* p = nullptr;
    explicit PinnedBuf(size_t bytes) { QK_HIP(hipHostMalloc(&p, std::max<size_t>(bytes, 16))); }
    ~PinnedBuf() { (void)hipHostFree(p); }
    template <typename T>
    T* get() const { return static_cast<T*>(p); }
};

}  // namespace

LanczosResult lanczos_run(const Hamiltonian& H, int maxiter, double tol, const cplx* start,
                          bool want_vector, const LanczosIO& io) {
    LanczosResult res;
    const u64 d = H.dim();
    if (d == 0) return res;
    hipStream_t st = H.stream;
    const int nmax = int(std::min<i64>(i64(maxiter), i64(d)));
    if (nmax <= 0) return res;
    constexpr int R = vec::LzStep::kRecFields;

    // The basis lives in slabs of kSlab vectors, allocated as the run needs
    // them; V (device) holds the pointer of every allocated vector.
    constexpr int kSlab = 8;
    std::vector<DevBuf> slabs;
    std::vector<cplx*> hptr(size_t(nmax), nullptr);   // persistent: source of async copies
    DevBuf ptrs(sizeof(void*) * size_t(nmax));
    auto vptr = [&](int i) -> cplx* {
        while (size_t(i / kSlab) >= slabs.size()) {
            const int first = int(slabs.size()) * kSlab;
            const int cnt = std::min(kSlab, nmax - first);
            slabs.emplace_back(sizeof(cplx) * d * size_t(cnt));
            for (int j = 0; j < cnt; ++j) hptr[size_t(first + j)] = slabs.back().get<cplx>() + u64(j) * d;
            QK_HIP(hipMemcpyAsync(ptrs.get<cplx*>() + first, &hptr[size_t(first)], sizeof(void*) * cnt,
                                  hipMemcpyHostToDevice, st));
        }
        return hptr[size_t(i)];
    };
    DevBuf w(sizeof(cplx) * d), red, sc(64 * sizeof(double));
    vec::LzStep lz(d, nmax);

    std::vector<double> alphas, betas;
    double e_old = 1e100;
    int first = 0;   // first step this call runs
    if (!io.resume_path.empty()) {
        LanczosState ck = load_lanczos_checkpoint(io.resume_path, H);
        if (ck.next_step > nmax) fail("lanczos resume: the checkpoint is past maxiter");
        first = ck.next_step;
        alphas = std::move(ck.alphas);
        betas = std::move(ck.betas);
        e_old = ck.e_old;
        for (int j = 0; j <= std::min(first, nmax - 1); ++j)
            QK_HIP(hipMemcpyAsync(vptr(j), ck.vectors.data() + u64(j) * d, sizeof(cplx) * d,
                                  hipMemcpyHostToDevice, st));
        QK_HIP(hipStreamSynchronize(st));
    } else {
        if (start) {
            QK_HIP(hipMemcpyAsync(vptr(0), start, sizeof(cplx) * d, hipMemcpyHostToDevice, st));
        } else {
            std::unique_ptr<cplx[]> s0(new cplx[d]);
            uniform_start(d, 1234, s0.get());
            QK_HIP(hipMemcpyAsync(vptr(0), s0.get(), sizeof(cplx) * d, hipMemcpyHostToDevice, st));
            QK_HIP(hipStreamSynchronize(st));
        }
        // normalize (vector_ops.hpp:91-97: unchanged when the norm is 0)
        const double nrm = dev_norm(vptr(0), d, red, sc, st);
        if (nrm != 0.0) vec::scale(vptr(0), d, cplx{1.0 / nrm, 0.0}, st);
    }
    lz.reset(u32(first), st);
    if (first > 0 && first <= nmax) {
        // the previous step's beta, read by the next step's three-term subtraction
        const double bprev = betas.empty() ? 0.0 : betas.back();
        QK_HIP(hipMemcpyAsync(lz.rec() + size_t(R) * (first - 1) + 4, &bprev, sizeof(double),
                              hipMemcpyHostToDevice, st));
        QK_HIP(hipStreamSynchronize(st));
    }

    // Short vectors: one step is one captured graph replayed with the step
    // index on the device (the apply reads a fixed copy of v_k); long
    // vectors: the same kernels launched directly, the apply reading v_k.
    const bool use_graph = d <= (u64(1) << 20) && std::getenv("QKRYLOV_LANCZOS_GRAPH") == nullptr
                               ? true
                               : (std::getenv("QKRYLOV_LANCZOS_GRAPH") != nullptr &&
                                  std::getenv("QKRYLOV_LANCZOS_GRAPH")[0] == '1');
    const bool fused = H.fused_dot(1);
    DevBuf apart(sizeof(double) * size_t(std::max(H.dot_parts(), 1)));
    DevBuf vcur;
    cplx* const* V = ptrs.get<cplx*>();
    auto step_body = [&](const cplx* x, cplx* vc, int coef_blocks) {
        if (fused) {
            H.apply_dot(x, w.get<cplx>(), apart.get<double>(), st);
            lz.alpha_from_parts(apart.get<double>(), u32(H.dot_parts()), st);
        } else {
            H.apply(x, w.get<cplx>(), 1, d, st);
            lz.alpha_dot(x, w.get<cplx>(), st);
        }
        lz.reorth(V, w.get<cplx>(), 1, coef_blocks, st);
        lz.reorth(V, w.get<cplx>(), 2, coef_blocks, st);
        lz.finish(V, w.get<cplx>(), vc, st);
    };
    StepGraph graph;
    if (use_graph && first < nmax) {
        vcur.resize(sizeof(cplx) * d);
        QK_HIP(hipMemcpyAsync(vcur.get(), vptr(first), sizeof(cplx) * d, hipMemcpyDeviceToDevice, st));
        QK_HIP(hipStreamSynchronize(st));
        graph.capture(st, [&] { step_body(vcur.get<cplx>(), vcur.get<cplx>(), nmax + 1); });
    }

    // Steps are queued ahead of the host's stopping test (the reference's
    // rule, lanczos.cpp:153-182, evaluated step by step on the records read
    // back), so the device never waits for the host; a step queued past the
    // stopping step is discarded (it changes nothing the result uses).
    PinnedBuf hrec(sizeof(double) * size_t(R) * size_t(nmax));
    struct Batch {
        int k0, k1;
        hipEvent_t ev;
    };
    std::vector<Batch> inflight;
    std::vector<hipEvent_t> evpool;
    auto get_event = [&]() {
        if (!evpool.empty()) {
            hipEvent_t e = evpool.back();
            evpool.pop_back();
            return e;
        }
        hipEvent_t e;
        QK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        return e;
    };
    auto cleanup_events = [&]() {
        for (auto& b : inflight) evpool.push_back(b.ev);
        inflight.clear();
        for (hipEvent_t e : evpool) (void)hipEventDestroy(e);
        evpool.clear();
    };
    const bool small = d <= (u64(1) << 20);
    const int batch = small ? 4 : 1;
    int enq = first;
    bool done = first >= nmax;
    double last_change = 1e100;
    res.iterations = first;
    try {
        while (!done) {
            const int evald = inflight.empty() ? enq : inflight.front().k0;
            int depth = small ? 2 * batch : (tol <= 0.0 || last_change > 1e3 * tol ? 2 : 1);
            if (tol <= 0.0) depth = std::max(depth, 16);
            while (enq < nmax && enq - evald < depth) {
                const int k1 = std::min(nmax, enq + batch);
                for (int k = enq; k < k1; ++k) {
                    if (k + 1 < nmax) vptr(k + 1);   // slot written by this step
                    if (use_graph) QK_HIP(hipGraphLaunch(graph.ge, st));
                    else step_body(vptr(k), nullptr, k + 2);
                }
                QK_HIP(hipMemcpyAsync(hrec.get<double>() + size_t(R) * enq, lz.rec() + size_t(R) * enq,
                                      sizeof(double) * size_t(R) * size_t(k1 - enq), hipMemcpyDeviceToHost,
                                      st));
                hipEvent_t ev = get_event();
                QK_HIP(hipEventRecord(ev, st));
                inflight.push_back(Batch{enq, k1, ev});
                enq = k1;
            }
            Batch b = inflight.front();
            inflight.erase(inflight.begin());
            QK_HIP(hipEventSynchronize(b.ev));
            evpool.push_back(b.ev);
            for (int iter = b.k0; iter < b.k1 && !done; ++iter) {
                const double* r = hrec.get<double>() + size_t(R) * iter;
                res.iterations = iter + 1;
                alphas.push_back(r[0]);
                if (r[5] != 0.0) ++res.second_passes;
                const double beta = r[4];
                if (beta < 1e-15) {
                    res.converged = true;
                    done = true;
                    break;
                }
                if (iter + 1 == nmax) {
                    done = true;
                    break;
                }
                betas.push_back(beta);
                if (iter > 0) {
                    const double e = tridiag_ground(alphas, betas, int(alphas.size()), nullptr);
                    last_change = std::abs(e - e_old);
                    if (last_change < tol) {
                        res.converged = true;
                        done = true;
                        break;
                    }
                    e_old = e;
                }
                const bool stop_here = io.stop_after > 0 && iter + 1 >= io.stop_after;
                if (!io.save_path.empty() &&
                    (stop_here || (io.save_every > 0 && (iter + 1) % io.save_every == 0))) {
                    // vectors v_0..v_{iter+1} are final once step iter is done
                    QK_HIP(hipStreamSynchronize(st));
                    LanczosState ck;
                    ck.next_step = iter + 1;
                    ck.alphas = alphas;
                    ck.betas = betas;
                    ck.e_old = e_old;
                    save_lanczos_checkpoint(io.save_path, H, ck, hptr.data(), st);
                }
                if (stop_here) {
                    done = true;
                    res.stopped = true;
                    break;
                }
            }
        }
        QK_HIP(hipStreamSynchronize(st));
    } catch (...) {
        (void)hipStreamSynchronize(st);
        cleanup_events();
        throw;
    }
    cleanup_events();

    std::vector<double> z;
    const int n = int(alphas.size());
    res.energy = tridiag_ground(alphas, betas, n, want_vector && !res.stopped ? &z : nullptr);
    if (want_vector && !res.stopped) {
        DevBuf ritz(sizeof(cplx) * d), coef(sizeof(cplx) * std::max(n, 1));
        QK_HIP(hipMemsetAsync(ritz.get(), 0, sizeof(cplx) * d, st));
        if (n > 0) {
            std::vector<cplx> cz(n);
            for (int i = 0; i < n; ++i) cz[i] = cplx{z[i], 0.0};
            QK_HIP(hipMemcpyAsync(coef.get(), cz.data(), sizeof(cplx) * n, hipMemcpyHostToDevice, st));
            vec::lincomb(V, n, coef.get<cplx>(), ritz.get<cplx>(), d, st);
            QK_HIP(hipStreamSynchronize(st));
        }
        const double nrm = dev_norm(ritz.get<cplx>(), d, red, sc, st);
        if (nrm != 0.0) vec::scale(ritz.get<cplx>(), d, cplx{1.0 / nrm, 0.0}, st);
        res.eigenvector.resize(d);
        QK_HIP(hipMemcpyAsync(res.eigenvector.data(), ritz.get(), sizeof(cplx) * d,
                              hipMemcpyDeviceToHost, st));
        QK_HIP(hipStreamSynchronize(st));
    }
    return res;
}

LanczosResult lanczos_ground_state(const Hamiltonian& H, int maxiter, double tol,
                                   const cplx* start, bool want_vector) {
    return lanczos_run(H, maxiter, tol, start, want_vector, LanczosIO{});
}

// ---------------------------------------------------------------------------
// Block Lanczos (no reference counterpart; extends lanczos.cpp:101-202)
// ---------------------------------------------------------------------------
namespace {

// Upper-triangular R with G = R^H R (G Hermitian positive definite, b x b,
// row-major); false when G is not numerically positive definite.
bool cholesky_upper(const std::vector<cplx>& G, int b, std::vector<cplx>& R) {
    R.assign(size_t(b) * b, cplx{0.0, 0.0});
    double gmax = 0.0;
    for (int i = 0; i < b; ++i) gmax = std::max(gmax, G[size_t(i) * b + i].re);
    for (int i = 0; i < b; ++i) {
        double s = G[size_t(i) * b + i].re;
        for (int k = 0; k < i; ++k) {
            const cplx r = R[size_t(k) * b + i];
            s -= r.re * r.re + r.im * r.im;
        }
        if (!(s > 1e-24 * std::max(gmax, 1e-300))) return false;
        const double rii = std::sqrt(s);
        R[size_t(i) * b + i] = cplx{rii, 0.0};
        for (int j = i + 1; j < b; ++j) {
            cplx t = G[size_t(i) * b + j];
            for (int k = 0; k < i; ++k) {   // t -= conj(R[k][i]) R[k][j]
                const cplx a = cconj(R[size_t(k) * b + i]), c = R[size_t(k) * b + j];
                t.re -= a.re * c.re - a.im * c.im;
                t.im -= a.re * c.im + a.im * c.re;
            }
            R[size_t(i) * b + j] = cplx{t.re / rii, t.im / rii};
        }
    }
    return true;
}

// inverse of an upper-triangular b x b matrix
std::vector<cplx> inv_upper(const std::vector<cplx>& R, int b) {
    std::vector<cplx> X(size_t(b) * b, cplx{0.0, 0.0});
    for (int j = 0; j < b; ++j) {
        // solve R x = e_j by back substitution
        for (int i = j; i >= 0; --i) {
            cplx s = (i == j) ? cplx{1.0, 0.0} : cplx{0.0, 0.0};
            for (int k = i + 1; k <= j; ++k) {
                const cplx a = R[size_t(i) * b + k], x = X[size_t(k) * b + j];
                s.re -= a.re * x.re - a.im * x.im;
                s.im -= a.re * x.im + a.im * x.re;
            }
            const cplx d = R[size_t(i) * b + i];   // real, positive
            X[size_t(i) * b + j] = cplx{s.re / d.re, s.im / d.re};
        }
    }
    return X;
}

std::vector<cplx> matmul(const std::vector<cplx>& A, const std::vector<cplx>& B, int n) {
    std::vector<cplx> C(size_t(n) * n, cplx{0.0, 0.0});
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            const cplx a = A[size_t(i) * n + k];
            for (int j = 0; j < n; ++j) {
                const cplx c = B[size_t(k) * n + j];
                C[size_t(i) * n + j].re += a.re * c.re - a.im * c.im;
                C[size_t(i) * n + j].im += a.re * c.im + a.im * c.re;
            }
        }
    return C;
}

}  // namespace

BlockLanczosResult block_lanczos_lowest(const Hamiltonian& H, int b, int nev, int maxiter,
                                        double tol, bool want_vectors) {
    BlockLanczosResult res;
    const u64 d = H.dim();
    if (d == 0) return res;
    b = std::max(1, std::min({b, 16, int(std::min<u64>(d, 16))}));
    nev = std::max(1, std::min(nev, b));
    const int nblk = std::max(1, std::min<int>(maxiter, int(d / u64(b))));
    hipStream_t st = H.stream;

    std::vector<DevBuf> Q;   // block j: b vectors of d, vector i at i*d
    DevBuf W(sizeof(cplx) * d * b), Wt(sizeof(cplx) * d * b), red, g, hcoef, coef;
    DevBuf qptr(sizeof(void*) * size_t(nblk + 1) * b), wptr(sizeof(void*) * b),
        wtptr(sizeof(void*) * b), lptr(sizeof(void*) * 2 * b);
    std::vector<const void*> qp;   // host copy of the pointer list
    auto set_ptrs = [&](DevBuf& arr, const std::vector<const void*>& p) {
        QK_HIP(hipMemcpyAsync(arr.get(), p.data(), sizeof(void*) * p.size(), hipMemcpyHostToDevice, st));
        QK_HIP(hipStreamSynchronize(st));
    };
    auto block_ptrs = [&](cplx* base) {
        std::vector<const void*> p(b);
        for (int i = 0; i < b; ++i) p[i] = base + u64(i) * d;
        return p;
    };
    set_ptrs(wptr, block_ptrs(W.get<cplx>()));
    set_ptrs(wtptr, block_ptrs(Wt.get<cplx>()));
    g.resize(sizeof(cplx) * 64 * 16);

    auto gram = [&](const cplx* const* A, int m, const cplx* const* B, int nb) {
        // <A_i, B_k> for i < m (in chunks of 64), k < nb
        std::vector<cplx> G(size_t(m) * nb);
        for (int c0 = 0; c0 < m; c0 += 64) {
            const int mc = std::min(64, m - c0);
            vec::block_dot(A + c0, mc, B, nb, d, g.get<cplx>(), red, st);
            QK_HIP(hipMemcpyAsync(G.data() + size_t(c0) * nb, g.get(), sizeof(cplx) * mc * nb,
                                  hipMemcpyDeviceToHost, st));
        }
        QK_HIP(hipStreamSynchronize(st));
        return G;
    };
    // W_k -= sum_i V_i C[i][k] for the m vectors V (device pointer list)
    auto subtract = [&](const cplx* const* V, int m, const std::vector<cplx>& C) {
        hcoef.resize(sizeof(cplx) * size_t(std::max(m, 1)) * b);
        std::vector<cplx> h(size_t(m) * b);
        for (int k = 0; k < b; ++k)
            for (int i = 0; i < m; ++i) h[size_t(k) * m + i] = C[size_t(i) * b + k];
        QK_HIP(hipMemcpyAsync(hcoef.get(), h.data(), sizeof(cplx) * h.size(), hipMemcpyHostToDevice, st));
        for (int k = 0; k < b; ++k)
            vec::msub(V, m, hcoef.get<cplx>() + size_t(k) * m, W.get<cplx>() + u64(k) * d, d, st);
    };
    // W = W R^{-1} with R from Cholesky-QR twice; false on rank loss
    auto cholqr = [&](std::vector<cplx>& Rtot) {
        Rtot.assign(size_t(b) * b, cplx{0.0, 0.0});
        for (int i = 0; i < b; ++i) Rtot[size_t(i) * b + i] = cplx{1.0, 0.0};
        for (int pass = 0; pass < 2; ++pass) {
            std::vector<cplx> G = gram(wptr.get<const cplx*>(), b, wptr.get<const cplx*>(), b);
            std::vector<cplx> R;
            if (!cholesky_upper(G, b, R)) return false;
            const std::vector<cplx> Ri = inv_upper(R, b);
            // out_k = sum_j Ri[j][k] W_j
            std::vector<cplx> c(size_t(b) * b);
            for (int k = 0; k < b; ++k)
                for (int j = 0; j < b; ++j) c[size_t(k) * b + j] = Ri[size_t(j) * b + k];
            coef.resize(sizeof(cplx) * c.size());
            QK_HIP(hipMemcpyAsync(coef.get(), c.data(), sizeof(cplx) * c.size(), hipMemcpyHostToDevice, st));
            vec::lincomb_many(wptr.get<const cplx*>(), b, coef.get<cplx>(), b, Wt.get<cplx>(), d, d, st);
            QK_HIP(hipMemcpyAsync(W.get(), Wt.get(), sizeof(cplx) * d * b, hipMemcpyDeviceToDevice, st));
            Rtot = matmul(R, Rtot, b);
        }
        return true;
    };
    auto push_block = [&]() {
        Q.emplace_back(sizeof(cplx) * d * b);
        QK_HIP(hipMemcpyAsync(Q.back().get(), W.get(), sizeof(cplx) * d * b, hipMemcpyDeviceToDevice, st));
        for (const void* p : block_ptrs(Q.back().get<cplx>())) qp.push_back(p);
        set_ptrs(qptr, qp);
    };

    // start block
    {
        std::unique_ptr<cplx[]> s0(new cplx[std::max<u64>(d * u64(b), 1)]);
        uniform_start(d * u64(b), 1234, s0.get());
        QK_HIP(hipMemcpyAsync(W.get(), s0.get(), sizeof(cplx) * d * b, hipMemcpyHostToDevice, st));
        QK_HIP(hipStreamSynchronize(st));
        std::vector<cplx> R0;
        if (!cholqr(R0)) fail("block Lanczos: start block is rank deficient");
        push_block();
    }

    std::vector<std::vector<cplx>> A, B;   // A_j, B_{j+1}
    std::vector<double> e_old;
    std::vector<double> ev;
    std::vector<cplx> S;
    int n = 0;
    for (int j = 0; j < nblk; ++j) {
        res.iterations = j + 1;
        const cplx* const* Qj = qptr.get<const cplx*>() + size_t(j) * b;
        H.apply(Q[j].get<cplx>(), W.get<cplx>(), b, d, st);
        std::vector<cplx> Aj = gram(Qj, b, wptr.get<const cplx*>(), b);
        for (int i = 0; i < b; ++i)            // Hermitian part
            for (int k = i; k < b; ++k) {
                const cplx x = Aj[size_t(i) * b + k], y = cconj(Aj[size_t(k) * b + i]);
                const cplx m2{0.5 * (x.re + y.re), 0.5 * (x.im + y.im)};
                Aj[size_t(i) * b + k] = m2;
                Aj[size_t(k) * b + i] = cconj(m2);
            }
        A.push_back(Aj);
        // three-term recurrence
        subtract(Qj, b, Aj);
        if (j > 0) {
            std::vector<cplx> BH(size_t(b) * b);   // B_j^H
            for (int i = 0; i < b; ++i)
                for (int k = 0; k < b; ++k) BH[size_t(i) * b + k] = cconj(B[j - 1][size_t(k) * b + i]);
            subtract(qptr.get<const cplx*>() + size_t(j - 1) * b, b, BH);
        }
        // full reorthogonalisation against every stored vector, twice
        const int m = (j + 1) * b;
        for (int pass = 0; pass < 2; ++pass) {
            std::vector<cplx> G = gram(qptr.get<const cplx*>(), m, wptr.get<const cplx*>(), b);
            for (int c0 = 0; c0 < m; c0 += 64) {
                const int mc = std::min(64, m - c0);
                std::vector<cplx> Gc(G.begin() + size_t(c0) * b, G.begin() + size_t(c0 + mc) * b);
                subtract(qptr.get<const cplx*>() + c0, mc, Gc);
            }
        }
        // eigenvalues of the block tridiagonal T (size m)
        n = m;
        const bool check = m <= 96 || (j % std::max(1, m / 96)) == 0 || j + 1 == nblk;
        if (check) {
            std::vector<cplx> T(size_t(m) * m, cplx{0.0, 0.0});
            for (int q = 0; q <= j; ++q) {
                for (int i = 0; i < b; ++i)
                    for (int k = 0; k < b; ++k)
                        T[size_t(q * b + i) * m + q * b + k] = A[q][size_t(i) * b + k];
                if (q < j)
                    for (int i = 0; i < b; ++i)
                        for (int k = 0; k < b; ++k) {
                            // T[q+1][q] = B_{q+1}, T[q][q+1] = B_{q+1}^H
                            const cplx x = B[q][size_t(i) * b + k];
                            T[size_t((q + 1) * b + i) * m + q * b + k] = x;
                            T[size_t(q * b + k) * m + (q + 1) * b + i] = cconj(x);
                        }
            }
            hermitian_eigh(T, m, ev, S);
            std::vector<double> e(ev.begin(), ev.begin() + std::min(nev, m));
            if (!e_old.empty() && e.size() == e_old.size()) {
                double dmax = 0.0;
                for (size_t i = 0; i < e.size(); ++i) dmax = std::max(dmax, std::abs(e[i] - e_old[i]));
                if (dmax < tol) {
                    res.converged = true;
                    break;
                }
            }
            e_old = e;
        }
        if (j + 1 == nblk) break;
        std::vector<cplx> Bn;
        if (!cholqr(Bn)) {   // the Krylov space is exhausted
            res.converged = true;
            break;
        }
        B.push_back(Bn);
        push_block();
    }
    // final Ritz pairs from the T of the last completed block
    {
        const int m = res.iterations * b;
        std::vector<cplx> T(size_t(m) * m, cplx{0.0, 0.0});
        for (int q = 0; q < res.iterations; ++q) {
            for (int i = 0; i < b; ++i)
                for (int k = 0; k < b; ++k) T[size_t(q * b + i) * m + q * b + k] = A[q][size_t(i) * b + k];
            if (q + 1 < res.iterations)
                for (int i = 0; i < b; ++i)
                    for (int k = 0; k < b; ++k) {
                        const cplx x = B[q][size_t(i) * b + k];
                        T[size_t((q + 1) * b + i) * m + q * b + k] = x;
                        T[size_t(q * b + k) * m + (q + 1) * b + i] = cconj(x);
                    }
        }
        hermitian_eigh(T, m, ev, S);
        n = m;
    }
    const int ne = std::min(nev, n);
    res.eigenvalues.assign(ev.begin(), ev.begin() + ne);
    if (want_vectors && ne > 0) {
        coef.resize(sizeof(cplx) * size_t(ne) * n);
        QK_HIP(hipMemcpyAsync(coef.get(), S.data(), sizeof(cplx) * size_t(ne) * n, hipMemcpyHostToDevice, st));
        DevBuf E(sizeof(cplx) * d * ne);
        vec::lincomb_many(qptr.get<const cplx*>(), n, coef.get<cplx>(), ne, E.get<cplx>(), d, d, st);
        res.eigenvectors.assign(ne, std::vector<cplx>(d));
        for (int k = 0; k < ne; ++k)
            QK_HIP(hipMemcpyAsync(res.eigenvectors[k].data(), E.get<cplx>() + u64(k) * d, sizeof(cplx) * d,
                                  hipMemcpyDeviceToHost, st));
        QK_HIP(hipStreamSynchronize(st));
    }
    return res;
}

// ---------------------------------------------------------------------------
// Davidson (reference src/solvers/davidson.cpp:85-227)
// ---------------------------------------------------------------------------
namespace {

// Cholesky factor of a k x k Hermitian Gram matrix G (row-major) with
// dependent columns dropped: R upper triangular over the kept columns,
// kept[j] = column j kept.  A column whose remaining squared norm is 1e-20
// or less is dropped; false when a kept pivot holds less than rel_min of its
// column's squared norm (the factorisation would be inaccurate).
bool gram_cholesky(const std::vector<cplx>& G, int k, std::vector<cplx>& R, std::vector<char>& kept,
                   double rel_min) {
    R.assign(size_t(k) * k, cplx{0.0, 0.0});
    kept.assign(k, 0);
    for (int j = 0; j < k; ++j) {
        double d2 = G[size_t(j) * k + j].re;
        for (int i = 0; i < j; ++i) {
            if (!kept[i]) continue;
            cplx sum = G[size_t(i) * k + j];
            for (int q = 0; q < i; ++q) {
                if (!kept[q]) continue;
                const cplx a = R[size_t(q) * k + i], b = R[size_t(q) * k + j];
                sum.re -= a.re * b.re + a.im * b.im;   // conj(a) b
                sum.im -= a.re * b.im - a.im * b.re;
            }
            const double rii = R[size_t(i) * k + i].re;
            R[size_t(i) * k + j] = cplx{sum.re / rii, sum.im / rii};
            d2 -= (sum.re * sum.re + sum.im * sum.im) / (rii * rii);
        }
        if (d2 <= 1e-20) continue;
        if (d2 < rel_min * G[size_t(j) * k + j].re) return false;
        R[size_t(j) * k + j] = cplx{std::sqrt(d2), 0.0};
        kept[j] = 1;
    }
    return true;
}

// Coefficients of Q = T R^-1 for the kept columns: c[o * k + i] for output o
// (the o-th kept column) and input i; returns the number of outputs.
int inverse_coefficients(const std::vector<cplx>& R, const std::vector<char>& kept, int k,
                         std::vector<cplx>& c) {
    std::vector<int> ids;
    for (int j = 0; j < k; ++j)
        if (kept[j]) ids.push_back(j);
    const int nk = int(ids.size());
    c.assign(size_t(std::max(nk, 1)) * k, cplx{0.0, 0.0});
    // column o of R^-1 over the kept indices, by back substitution
    for (int o = 0; o < nk; ++o) {
        const int jo = ids[o];
        std::vector<cplx> x(k, cplx{0.0, 0.0});
        const double rjj = R[size_t(jo) * k + jo].re;
        x[jo] = cplx{1.0 / rjj, 0.0};
        for (int p = o - 1; p >= 0; --p) {
            const int i = ids[p];
            cplx sum{0.0, 0.0};
            for (int q = p + 1; q <= o; ++q) {
                const int l = ids[q];
                const cplx a = R[size_t(i) * k + l], b = x[l];
                sum.re += a.re * b.re - a.im * b.im;
                sum.im += a.re * b.im + a.im * b.re;
            }
            const double rii = R[size_t(i) * k + i].re;
            x[i] = cplx{-sum.re / rii, -sum.im / rii};
        }
        for (int i = 0; i < k; ++i) c[size_t(o) * k + i] = x[i];
    }
    return nk;
}

}  // namespace

// Two Cholesky-QR passes over the k vectors of T (pointer table pT): the
// first writes Q1 to `tmp` (table pTmp), the second writes the orthonormal
// result to `out` (consecutive vectors of n).  Returns the number of vectors
// written, or -1 (nothing written to `out`) when a pivot is too small for
// the Gram route.
static int cholqr2_into(const cplx* const* pT, const cplx* const* pTmp, int k, cplx* tmp, cplx* out,
                        u64 n, DevBuf& h, DevBuf& coef, DevBuf& red, hipStream_t st) {
    std::vector<cplx> G, R, c;
    std::vector<char> kept;
    const cplx* const* src = pT;
    int cur = k;
    for (int pass = 0; pass < 2; ++pass) {
        h.resize(sizeof(cplx) * size_t(cur) * cur);
        vec::block_dot(src, cur, src, cur, n, h.get<cplx>(), red, st);
        G.resize(size_t(cur) * cur);
        QK_HIP(hipMemcpyAsync(G.data(), h.get(), sizeof(cplx) * G.size(), hipMemcpyDeviceToHost, st));
        QK_HIP(hipStreamSynchronize(st));
        if (!gram_cholesky(G, cur, R, kept, 1e-6)) return -1;
        const int nk = inverse_coefficients(R, kept, cur, c);
        if (nk == 0) return 0;
        coef.resize(sizeof(cplx) * c.size());
        QK_HIP(hipMemcpyAsync(coef.get(), c.data(), sizeof(cplx) * c.size(), hipMemcpyHostToDevice, st));
        vec::lincomb_many(src, cur, coef.get<cplx>(), nk, pass == 0 ? tmp : out, n, n, st);
        QK_HIP(hipStreamSynchronize(st));   // c is reused by the next pass
        src = pTmp;
        cur = nk;
    }
    return cur;
}

DavidsonResult davidson_lowest(const Hamiltonian& H, int n_eig, int max_subspace, double tol,
                               bool want_vectors, Preconditioner* pre, int max_iter) {
    DavidsonResult res;
    const u64 d = H.dim();
    if (d == 0) return res;
    hipStream_t st = H.stream;
    n_eig = int(std::min<i64>(n_eig, i64(d)));
    max_subspace = std::max(max_subspace, 2 * n_eig);
    // the basis never exceeds max_subspace vectors: a step that would pass it
    // restarts to n_eig first and then adds at most n_eig (:196-223)
    const int cap = max_subspace + 1;
    const int nb = std::max(n_eig, 1);

    // All device storage is allocated once: the basis V and H V as slabs of
    // cap vectors with fixed pointer tables, and n_eig-vector slabs for the
    // residuals, Ritz vectors, corrections and the restart.
    DevBuf diag(sizeof(cplx) * d), red, sc(64 * sizeof(double)), h(sizeof(cplx) * cap * cap),
        slabV(sizeof(cplx) * d * cap), slabHV(sizeof(cplx) * d * cap), Rb(sizeof(cplx) * d * nb),
        Ub(sizeof(cplx) * d * nb), Tb(sizeof(cplx) * d * nb), Nb(sizeof(cplx) * d * nb * 2),
        nrm(sizeof(double) * nb), coef(sizeof(cplx) * size_t(nb) * 2 * cap),
        tnorm(sizeof(double) * 3 * nb);
    auto vp = [&](int i) { return slabV.get<cplx>() + u64(i) * d; };
    auto hvp = [&](int i) { return slabHV.get<cplx>() + u64(i) * d; };
    DevBuf pV(sizeof(void*) * cap), pHV(sizeof(void*) * cap), pI(sizeof(void*) * 2 * cap),
        pT(sizeof(void*) * nb), pN(sizeof(void*) * nb);
    {
        std::vector<const void*> a(cap), b(cap), c(2 * size_t(cap)), t(nb), nn(nb);
        for (int j = 0; j < nb; ++j) t[j] = Tb.get<cplx>() + u64(j) * d;
        for (int j = 0; j < nb; ++j) nn[j] = Nb.get<cplx>() + u64(j) * d;
        QK_HIP(hipMemcpyAsync(pT.get(), t.data(), sizeof(void*) * nb, hipMemcpyHostToDevice, st));
        QK_HIP(hipMemcpyAsync(pN.get(), nn.data(), sizeof(void*) * nb, hipMemcpyHostToDevice, st));
        for (int i = 0; i < cap; ++i) {
            a[i] = vp(i);
            b[i] = hvp(i);
            c[2 * size_t(i)] = hvp(i);   // residual operands interleaved: HV_i, V_i
            c[2 * size_t(i) + 1] = vp(i);
        }
        QK_HIP(hipMemcpyAsync(pV.get(), a.data(), sizeof(void*) * cap, hipMemcpyHostToDevice, st));
        QK_HIP(hipMemcpyAsync(pHV.get(), b.data(), sizeof(void*) * cap, hipMemcpyHostToDevice, st));
        QK_HIP(hipMemcpyAsync(pI.get(), c.data(), sizeof(void*) * c.size(), hipMemcpyHostToDevice, st));
    }

    H.diagonal(diag.get<cplx>(), st);
    std::vector<cplx> dh(d);
    QK_HIP(hipMemcpyAsync(dh.data(), diag.get(), sizeof(cplx) * d, hipMemcpyDeviceToHost, st));
    QK_HIP(hipStreamSynchronize(st));
    std::vector<u64> idx(d);
    std::iota(idx.begin(), idx.end(), u64(0));
    std::sort(idx.begin(), idx.end(), [&](u64 i, u64 j) { return dh[i].re < dh[j].re; });

    // Start vectors: unit vectors on the n_eig smallest Re(diag) (:111-139).
    // Distinct unit vectors are exactly orthonormal, so the reference's
    // orthogonalisation leaves them unchanged and its random fallback never
    // triggers.
    QK_HIP(hipMemsetAsync(vp(0), 0, sizeof(cplx) * d * n_eig, st));
    const cplx one{1.0, 0.0};
    for (int i = 0; i < n_eig; ++i)
        QK_HIP(hipMemcpyAsync(vp(i) + idx[i], &one, sizeof(cplx), hipMemcpyHostToDevice, st));
    H.apply(vp(0), hvp(0), n_eig, d, st);
    QK_HIP(hipStreamSynchronize(st));
    int m = n_eig;

    // the first ne columns of S (Ritz coefficients), uploaded to coef
    auto upload_ritz = [&](const std::vector<cplx>& S, int mm, int ne) {
        QK_HIP(hipMemcpyAsync(coef.get(), S.data(), sizeof(cplx) * size_t(ne) * mm, hipMemcpyHostToDevice, st));
    };

    std::vector<cplx> Hs;   // projected matrix, row-major m*m
    int valid_rows = 0;
    max_iter = std::max(max_iter, 1);
    const bool use_block_dot = [] {
        const char* e = std::getenv("QKRYLOV_DAVIDSON_BLOCK_DOT");
        return !(e && std::string(e) == "0");
    }();
    PhaseClock clk("davidson");
    for (int iter = 0; iter < max_iter; ++iter) {
        clk.mark(st, "apply (expansion)");
        // H_sub[i][j] = <V_i, HV_j>, j <= i (:143-150); new rows only.
        std::vector<cplx> Hn(size_t(m) * m, cplx{0.0, 0.0});
        for (int i = 0; i < std::min(valid_rows, m); ++i)
            for (int j = 0; j < std::min(valid_rows, m); ++j) Hn[size_t(i) * m + j] = Hs[size_t(i) * valid_rows + j];
        const int nnew = m - valid_rows;
        if (use_block_dot && nnew > 0 && m <= 64 && nnew <= 16) {
            // one FP64-MFMA projection G = HV^H V_new for all new rows
            vec::block_dot(pHV.get<const cplx*>(), m, pV.get<const cplx*>() + valid_rows, nnew, d,
                           h.get<cplx>(), red, st);
            std::vector<cplx> G(size_t(m) * nnew);
            QK_HIP(hipMemcpyAsync(G.data(), h.get(), sizeof(cplx) * G.size(), hipMemcpyDeviceToHost, st));
            QK_HIP(hipStreamSynchronize(st));
            for (int i = valid_rows; i < m; ++i)
                for (int j = 0; j <= i; ++j) {
                    const cplx vij = cconj(G[size_t(j) * nnew + (i - valid_rows)]);   // <V_i, HV_j>
                    Hn[size_t(i) * m + j] = vij;
                    if (i != j) Hn[size_t(j) * m + i] = cconj(vij);
                }
        } else {
            for (int i = valid_rows; i < m; ++i) {
                vec::mdot(pHV.get<const cplx*>(), i + 1, vp(i), d, h.get<cplx>(), red, st);
                std::vector<cplx> hv(i + 1);
                QK_HIP(hipMemcpyAsync(hv.data(), h.get(), sizeof(cplx) * (i + 1), hipMemcpyDeviceToHost, st));
                QK_HIP(hipStreamSynchronize(st));
                for (int j = 0; j <= i; ++j) {
                    const cplx vij = cconj(hv[j]);   // <V_i, HV_j>
                    Hn[size_t(i) * m + j] = vij;
                    if (i != j) Hn[size_t(j) * m + i] = cconj(vij);
                }
            }
        }
        Hs = std::move(Hn);
        valid_rows = m;

        clk.mark(st, "projection");
        std::vector<double> ev;
        std::vector<cplx> S;
        hermitian_eigh(Hs, m, ev, S);
        clk.mark(st, "small eigensolve");

        // residuals r_k = sum_i s_ki (HV_i - lambda_k V_i) (:161-165)
        const int ne = std::min(n_eig, m);
        {
            std::vector<cplx> C(size_t(ne) * 2 * m);
            for (int k = 0; k < ne; ++k)
                for (int i = 0; i < m; ++i) {
                    const cplx s = S[size_t(k) * m + i];
                    C[size_t(k) * 2 * m + 2 * i] = s;
                    C[size_t(k) * 2 * m + 2 * i + 1] = cplx{-ev[k] * s.re, -ev[k] * s.im};
                }
            QK_HIP(hipMemcpyAsync(coef.get(), C.data(), sizeof(cplx) * C.size(), hipMemcpyHostToDevice, st));
            vec::lincomb_many(pI.get<const cplx*>(), 2 * m, coef.get<cplx>(), ne, Rb.get<cplx>(), d, d, st);
        }
        vec::norms(Rb.get<cplx>(), ne, d, d, nrm.get<double>(), red, st);
        std::vector<double> rn(ne);
        QK_HIP(hipMemcpyAsync(rn.data(), nrm.get(), sizeof(double) * ne, hipMemcpyDeviceToHost, st));
        QK_HIP(hipStreamSynchronize(st));
        clk.mark(st, "residuals");
        res.residuals.assign(ne, 0.0);
        bool all_conv = true;
        std::vector<int> need;
        for (int k = 0; k < ne; ++k) {
            res.residuals[k] = std::sqrt(rn[k]);
            if (res.residuals[k] > tol) {
                all_conv = false;
                need.push_back(k);
            }
        }
        if (all_conv || iter == max_iter - 1) {
            res.iterations = iter + 1;
            res.converged = all_conv;
            // Eigenvalues as Rayleigh quotients <u, Hu>/<u, u> of the Ritz
            // vectors (Hu = HV s): the Ritz values themselves while V is
            // orthonormal, and still second order in the residual where
            // rounding has cost V some orthogonality.
            std::vector<int> order(ne);
            std::iota(order.begin(), order.end(), 0);
            if (ne > 0) {
                upload_ritz(S, m, ne);
                cplx* hu = Nb.get<cplx>();
                vec::lincomb_many(pV.get<const cplx*>(), m, coef.get<cplx>(), ne, Ub.get<cplx>(), d, d, st);
                vec::lincomb_many(pHV.get<const cplx*>(), m, coef.get<cplx>(), ne, hu, d, d, st);
                for (int k = 0; k < ne; ++k)
                    vec::dot(Ub.get<cplx>() + u64(k) * d, hu + u64(k) * d, d, h.get<cplx>() + k, red, st);
                vec::norms(Ub.get<cplx>(), ne, d, d, nrm.get<double>(), red, st);
                std::vector<cplx> uhu(ne);
                std::vector<double> uu(ne);
                QK_HIP(hipMemcpyAsync(uhu.data(), h.get(), sizeof(cplx) * ne, hipMemcpyDeviceToHost, st));
                QK_HIP(hipMemcpyAsync(uu.data(), nrm.get(), sizeof(double) * ne, hipMemcpyDeviceToHost, st));
                QK_HIP(hipStreamSynchronize(st));
                for (int k = 0; k < ne; ++k)
                    if (uu[k] > 0.0) ev[k] = uhu[k].re / uu[k];
                std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ev[a] < ev[b]; });
            }
            for (int k = 0; k < ne; ++k) res.eigenvalues.push_back(ev[order[k]]);
            std::vector<double> rs(ne);
            for (int k = 0; k < ne; ++k) rs[k] = res.residuals[order[k]];
            res.residuals = rs;
            if (want_vectors && ne > 0) {
                res.eigenvectors.assign(ne, std::vector<cplx>(d));
                for (int k = 0; k < ne; ++k)
                    QK_HIP(hipMemcpyAsync(res.eigenvectors[k].data(), Ub.get<cplx>() + u64(order[k]) * d,
                                          sizeof(cplx) * d, hipMemcpyDeviceToHost, st));
                QK_HIP(hipStreamSynchronize(st));
            }
            return res;
        }

        // corrections t = M(r) (:166-177)
        const bool ritz = pre && pre->wants_ritz() && ne > 0;
        if (ritz) {
            upload_ritz(S, m, ne);
            vec::lincomb_many(pV.get<const cplx*>(), m, coef.get<cplx>(), ne, Ub.get<cplx>(), d, d, st);
        }
        const int nt = int(need.size());
        if (pre) {
            std::vector<const cplx*> rp(nt), up(nt);
            std::vector<cplx*> tp(nt);
            std::vector<double> lam(nt);
            for (int j = 0; j < nt; ++j) {
                rp[j] = Rb.get<cplx>() + u64(need[j]) * d;
                up[j] = ritz ? Ub.get<cplx>() + u64(need[j]) * d : nullptr;
                tp[j] = Tb.get<cplx>() + u64(j) * d;
                lam[j] = ev[need[j]];
            }
            pre->apply_many(H, nt, rp.data(), diag.get<cplx>(), lam.data(), ritz ? up.data() : nullptr,
                            tp.data(), st);
        } else {
            for (int j = 0; j < nt; ++j)
                vec::jacobi_precond(Rb.get<cplx>() + u64(need[j]) * d, diag.get<cplx>(), ev[need[j]],
                                    Tb.get<cplx>() + u64(j) * d, d, st);
        }

        clk.mark(st, "preconditioner");
        // restart to the n_eig Ritz pairs (:196-211)
        if (m + nt > max_subspace) {
            upload_ritz(S, m, ne);
            cplx* nv = Nb.get<cplx>();
            cplx* nhv = nv + u64(ne) * d;
            vec::lincomb_many(pV.get<const cplx*>(), m, coef.get<cplx>(), ne, nv, d, d, st);
            vec::lincomb_many(pHV.get<const cplx*>(), m, coef.get<cplx>(), ne, nhv, d, d, st);
            QK_HIP(hipMemcpyAsync(vp(0), nv, sizeof(cplx) * d * ne, hipMemcpyDeviceToDevice, st));
            QK_HIP(hipMemcpyAsync(hvp(0), nhv, sizeof(cplx) * d * ne, hipMemcpyDeviceToDevice, st));
            m = ne;
            valid_rows = 0;
        }

        // orthogonalise and expand (:213-223).  The reference orthogonalises
        // each correction against the basis including the corrections accepted
        // before it; here all corrections are first orthogonalised against the
        // old basis together (two block passes, each basis vector read once),
        // then one by one against the accepted ones, which spans the same
        // space.  H is applied to the accepted vectors together, after the
        // loop (only V enters the orthogonalisation).
        const int m0 = m;
        const bool block = use_block_dot && nt > 1 && m0 <= 64 && nt <= 16;
        double* tn2 = tnorm.get<double>();   // [0, nt): before, [nt, 2nt): after the basis
        cplx* T = Tb.get<cplx>();
        vec::norms(T, nt, d, d, tn2, red, st);
        std::vector<double> tn(2 * size_t(nt));
        auto fetch_norms = [&](int first, int count) {
            vec::norms(T, nt, d, d, tn2 + nt, red, st);
            QK_HIP(hipMemcpyAsync(tn.data() + first, tn2 + first, sizeof(double) * count,
                                  hipMemcpyDeviceToHost, st));
            QK_HIP(hipStreamSynchronize(st));
        };
        auto block_pass = [&]() {
            vec::block_dot(pV.get<const cplx*>(), m0, pT.get<const cplx*>(), nt, d, h.get<cplx>(), red, st);
            vec::block_sub(pV.get<const cplx*>(), m0, pT.get<cplx*>(), nt, h.get<cplx>(), d, st);
        };
        if (block) {
            // one block projection, and a second only when a correction lost
            // more than half its squared norm in the first (the Daniel-Gragg-
            // Kaufman-Stewart test, as in Lanczos)
            block_pass();
            if (nt > 0) fetch_norms(0, 2 * nt);
            bool again = false;
            for (int j = 0; j < nt; ++j) again = again || tn[nt + j] < 0.5 * tn[j];
            if (again) {
                block_pass();
                fetch_norms(nt, nt);
            }
        } else {
            for (int j = 0; j < nt; ++j) cgs2(pV, m0, T + u64(j) * d, d, h, red, st);
            if (nt > 0) fetch_norms(0, 2 * nt);
        }
        // a correction that lost most of its norm to the basis is left
        // orthogonal to it only to rounding relative to its original norm:
        // one more pass
        for (int j = 0; j < nt; ++j)
            if (std::sqrt(tn[nt + j]) > 1e-10 && std::sqrt(tn[nt + j]) < 1e-3 * std::sqrt(tn[j])) {
                cgs2(pV, m0, T + u64(j) * d, d, h, red, st);
                ++clk.extra;
            }
        // Among the corrections: Cholesky QR twice (Gram matrix, host
        // factorisation, Q = T R^-1), the second pass writing straight into
        // the basis slots.  A correction whose remaining norm after the
        // earlier ones is 1e-10 or less is dropped, as in sequential
        // Gram-Schmidt.  When a pivot keeps less than 1e-6 of its column's
        // squared norm the Gram route would lose accuracy, and the
        // corrections go through sequential CGS2 instead.
        const int mq = block && nt > 1 ? cholqr2_into(pT.get<const cplx*>(), pN.get<const cplx*>(), nt,
                                                      Nb.get<cplx>(), vp(m), d, h, coef, red, st)
                                       : -1;
        if (mq >= 0) {
            m += mq;
        } else {
            // in order and without host round trips: each is orthogonalised
            // against the earlier ones and normalised, or zeroed when its norm
            // is 1e-10 or less (a zero vector adds nothing to the later
            // projections, as a dropped one would not)
            double* inv = tn2 + 2 * nt;
            for (int j = 0; j < nt; ++j) {
                cplx* t = T + u64(j) * d;
                if (j > 0) cgs2(pT.get<const cplx*>(), j, t, d, h, red, st);
                vec::norm2(t, d, tn2 + j, red, st);
                vec::inv_norm_or_zero(tn2 + j, inv + j, 1, 1e-10, st);
                vec::scale_to(t, t, d, inv + j, st);
            }
            if (nt > 0) {
                QK_HIP(hipMemcpyAsync(tn.data(), tn2, sizeof(double) * nt, hipMemcpyDeviceToHost, st));
                QK_HIP(hipStreamSynchronize(st));
            }
            for (int j = 0; j < nt; ++j)
                if (std::sqrt(tn[j]) > 1e-10) {
                    QK_HIP(hipMemcpyAsync(vp(m), T + u64(j) * d, sizeof(cplx) * d, hipMemcpyDeviceToDevice, st));
                    ++m;
                }
        }
        clk.mark(st, "restart + orthogonalise");
        if (m > m0) H.apply(vp(m0), hvp(m0), m - m0, d, st);
    }
    return DavidsonResult{};   // not reached (the last iteration returns), as :226
}

// ---------------------------------------------------------------------------
// Block three-term recurrence without reorthogonalisation, shared by FTLM
// (reference ftlm.cpp:24-53) and continued fractions (dynamics.cpp:12-59).
// Vectors are kept unnormalised: v_k = c_k u_k, so one step is one apply
// plus one fused update pass.
// ---------------------------------------------------------------------------
namespace {

struct BlockRec {
    struct Grp {
        const Hamiltonian* h;
        int nvec;
        u64 off;
    };
    std::vector<Grp> grps;
    Segments seg;
    DevBuf bufU, bufY, bufP, c, cp, bp, alive, alpha, beta, nrm;
    DevBuf dparts, dcount;   // alpha partials written by the apply kernel
    u32 dstride = 0;         // 0: alpha from a separate pass over U and Y
    bool mparts = false;     // partials from the multi-sector launch (per tile)
    int nsamp = 0;
    hipStream_t st = nullptr;
    MultiApply multi;   // all groups in one apply launch when they allow it

    // allow_fused: alpha may come from the apply kernel's partial sums (else
    // always from the separate pass, whose per-sample reduction is the same
    // whatever other samples share the recurrence); multi_single: a single
    // vector may take the multi-sector launch (else the single-vector kernel)
    void setup(const std::vector<std::pair<const Hamiltonian*, int>>& g, hipStream_t stream,
               bool allow_fused = true, bool multi_single = true) {
        st = stream;
        std::vector<u64> lens;
        u64 off = 0;
        grps.clear();
        for (auto& x : g) {
            grps.push_back({x.first, x.second, off});
            for (int r = 0; r < x.second; ++r) lens.push_back(x.first->dim());
            off += x.first->dim() * u64(x.second);
        }
        seg.build(lens, st);
        nsamp = seg.nsamp;
        // One launch for every group and vector pays off while launches, not
        // bytes, dominate (small sectors); large bases keep the per-vector
        // kernels, which the block kernel does not beat there.
        std::vector<MultiApply::Group> mg;
        u64 dmax = 0;
        int ntot = 0;
        for (const Grp& x : grps) {
            mg.push_back({x.h, x.nvec, x.off});
            dmax = std::max(dmax, x.h->dim());
            ntot += x.nvec;
        }
        // A single vector takes the multi-sector launch too when alpha may
        // come from its tile partials: the same tiles and sums as in a
        // batch, so one vector alone and the batch agree bit for bit.
        const char* e = std::getenv("QKRYLOV_MULTI_APPLY");
        const char* f = std::getenv("QKRYLOV_FUSED_ALPHA");
        const bool fuse_ok = allow_fused && !(f && std::string(f) == "0");
        if (!(e && std::string(e) == "0") && (ntot > 1 || (fuse_ok && multi_single)) && ntot >= 1 &&
            dmax < (u64(1) << 22))
            multi.build(mg, st);
        // The apply kernels can return alpha's partial sums themselves (they
        // hold x and y when they store), which saves reading both vectors
        // again every step: per tile from the multi-sector launch, per
        // workgroup from per-vector single-vector launches.
        dstride = 0;
        mparts = false;
        if (fuse_ok && multi.tile_parts() && multi.unit_tiles().size() == size_t(nsamp)) {
            mparts = true;
            dstride = std::max<u32>(multi.part_stride(), 1);
            const size_t pb = sizeof(double) * size_t(dstride) * std::max<size_t>(nsamp, 1);
            dparts.resize(pb);
            QK_HIP(hipMemsetAsync(dparts.get(), 0, pb, st));
            upload(dcount, multi.unit_tiles().data(), multi.unit_tiles().size(), st);
        }
        bool fused = fuse_ok && !multi.ok() && !grps.empty();
        for (const Grp& x : grps)
            if (!(x.h->fused_dot(x.nvec) || x.h->dim() == 0)) fused = false;
        if (fused) {
            std::vector<u32> cnt;
            for (const Grp& x : grps) {
                dstride = std::max<u32>(dstride, u32(x.h->dot_parts()));
                for (int r = 0; r < x.nvec; ++r) cnt.push_back(x.h->dim() ? u32(x.h->dot_parts()) : 0u);
            }
            dstride = std::max<u32>(dstride, 1);
            const size_t pb = sizeof(double) * size_t(dstride) * std::max<size_t>(cnt.size(), 1);
            dparts.resize(pb);
            QK_HIP(hipMemsetAsync(dparts.get(), 0, pb, st));
            upload(dcount, cnt.data(), cnt.size(), st);
        }
        const size_t bytes = sizeof(cplx) * std::max<u64>(seg.total, 1);
        bufU.resize(bytes);
        bufY.resize(bytes);
        bufP.resize(bytes);
        const size_t ns = std::max(nsamp, 1);
        c.resize(sizeof(double) * ns);
        cp.resize(sizeof(double) * ns);
        bp.resize(sizeof(double) * ns);
        alive.resize(sizeof(int) * ns);
        nrm.resize(sizeof(double) * ns);
    }

    // U holds the start vectors; returns their norms (host).
    std::vector<double> start(double thresh) {
        std::vector<double> n(nsamp);
        vec::seg_norm(seg, bufU.get<cplx>(), nrm.get<double>(), st);
        QK_HIP(hipMemcpyAsync(n.data(), nrm.get(), sizeof(double) * nsamp, hipMemcpyDeviceToHost, st));
        QK_HIP(hipStreamSynchronize(st));
        std::vector<double> cv(nsamp), zero(nsamp, 0.0);
        std::vector<int> al(nsamp);
        for (int j = 0; j < nsamp; ++j) {
            al[j] = n[j] >= thresh ? 1 : 0;
            // normalize() leaves a zero vector unchanged (vector_ops.hpp:94-96)
            cv[j] = n[j] > 0.0 ? 1.0 / n[j] : 1.0;
        }
        QK_HIP(hipMemcpyAsync(c.get(), cv.data(), sizeof(double) * nsamp, hipMemcpyHostToDevice, st));
        QK_HIP(hipMemcpyAsync(cp.get(), zero.data(), sizeof(double) * nsamp, hipMemcpyHostToDevice, st));
        QK_HIP(hipMemcpyAsync(bp.get(), zero.data(), sizeof(double) * nsamp, hipMemcpyHostToDevice, st));
        QK_HIP(hipMemcpyAsync(alive.get(), al.data(), sizeof(int) * nsamp, hipMemcpyHostToDevice, st));
        QK_HIP(hipStreamSynchronize(st));
        return n;
    }

    // M steps; alpha/beta on the host as [step][sample].  With `graph` the
    // M steps (every apply and recurrence kernel, no host work in between:
    // alpha, beta and the scales stay on the device) are captured once into
    // a hipGraph and launched as one unit.  run() = prepare, launch, collect;
    // split, they let the host work on one recurrence's results while the
    // GPU runs another's steps.
    void run(int M, std::vector<double>& ah, std::vector<double>& bh, bool graph = false) {
        prepare(M, graph);
        launch();
        collect(ah, bh, st);
    }

    // alpha/beta cleared; with `graph`, the steps captured (the stream idle)
    void prepare(int M, bool graph) {
        steps_ = M;
        const size_t tot = size_t(std::max(M, 1)) * std::max(nsamp, 1);
        alpha.resize(sizeof(double) * tot);
        beta.resize(sizeof(double) * tot);
        QK_HIP(hipMemsetAsync(alpha.get(), 0, sizeof(double) * tot, st));
        QK_HIP(hipMemsetAsync(beta.get(), 0, sizeof(double) * tot, st));
        release_graph();
        if (graph && M > 0) {
            QK_HIP(hipStreamSynchronize(st));
            QK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                steps(M);
            } catch (...) {
                (void)hipStreamEndCapture(st, &graph_);
                release_graph();
                throw;
            }
            QK_HIP(hipStreamEndCapture(st, &graph_));
            const hipError_t e = hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                exec_ = nullptr;
                release_graph();
                QK_HIP(e);
            }
        }
    }

    // the steps queued on the stream (the graph, or direct launches)
    void launch() {
        if (exec_) QK_HIP(hipGraphLaunch(exec_, st));
        else steps(steps_);
        if (!done_) QK_HIP(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
        QK_HIP(hipEventRecord(done_, st));
    }

    // waits for the steps and copies alpha/beta on stream cs (which may be
    // another stream than the one the steps run on)
    void collect(std::vector<double>& ah, std::vector<double>& bh, hipStream_t cs) {
        const int M = steps_;
        ah.resize(size_t(M) * nsamp);
        bh.resize(size_t(M) * nsamp);
        QK_HIP(hipStreamWaitEvent(cs, done_, 0));
        if (M > 0 && nsamp > 0) {
            QK_HIP(hipMemcpyAsync(ah.data(), alpha.get(), sizeof(double) * M * nsamp, hipMemcpyDeviceToHost, cs));
            QK_HIP(hipMemcpyAsync(bh.data(), beta.get(), sizeof(double) * M * nsamp, hipMemcpyDeviceToHost, cs));
        }
        QK_HIP(hipStreamSynchronize(cs));
        release_graph();
    }

    BlockRec() = default;
    BlockRec(const BlockRec&) = delete;
    BlockRec& operator=(const BlockRec&) = delete;
    ~BlockRec() {
        if (st && (exec_ || done_)) (void)hipStreamSynchronize(st);
        release_graph();
        if (done_) (void)hipEventDestroy(done_);
    }

private:
    int steps_ = 0;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    hipEvent_t done_ = nullptr;

    void release_graph() {
        if (exec_) (void)hipGraphExecDestroy(exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        exec_ = nullptr;
        graph_ = nullptr;
    }

    void steps(int M) {
        cplx* U = bufU.get<cplx>();
        cplx* Y = bufY.get<cplx>();
        cplx* P = bufP.get<cplx>();
        for (int step = 0; step < M; ++step) {
            double* a = alpha.get<double>() + size_t(step) * nsamp;
            double* b = beta.get<double>() + size_t(step) * nsamp;
            if (dstride) {
                if (mparts) {
                    multi.launch(U, Y, st, dparts.get<double>());
                } else {
                    size_t j = 0;
                    for (auto& g : grps)
                        for (int r = 0; r < g.nvec; ++r, ++j) {
                            const u64 o = g.off + u64(r) * g.h->dim();
                            double* part = dparts.get<double>() + j * dstride;
                            if (g.h->dim() && !g.h->apply_dot(U + o, Y + o, part, st))
                                fail("block recurrence: fused apply not available");
                        }
                }
                vec::alpha_from_parts(dparts.get<double>(), dstride, dcount.get<u32>(), nsamp,
                                      c.get<double>(), a, alive.get<int>(), st);
            } else {
                if (multi.ok()) multi.launch(U, Y, st);
                else
                    for (auto& g : grps) g.h->apply(U + g.off, Y + g.off, g.nvec, g.h->dim(), st);
                vec::seg_alpha(seg, U, Y, c.get<double>(), a, alive.get<int>(), st);
            }
            vec::seg_update(seg, Y, U, P, c.get<double>(), cp.get<double>(), bp.get<double>(), a,
                            b, alive.get<int>(), step == 0 ? 1 : 0, st);
            cplx* oldP = P;
            P = U;
            U = Y;
            Y = oldP;
        }
    }
};

// Sequence of sample j: alphas until (and including) the breakdown step,
// betas strictly before it (ftlm.cpp:44-46, dynamics.cpp:274-277).
void extract(const std::vector<double>& ah, const std::vector<double>& bh, int M, int ns, int j,
             std::vector<double>& a, std::vector<double>& b) {
    a.clear();
    b.clear();
    for (int k = 0; k < M; ++k) {
        a.push_back(ah[size_t(k) * ns + j]);
        const double bt = bh[size_t(k) * ns + j];
        if (bt < 1e-15) break;
        b.push_back(bt);
    }
}

}  // namespace

std::vector<CFResult> continued_fraction_block_dev(const Hamiltonian& H, const cplx* phi0_dev,
                                                   int nblock, int n_iter) {
    std::vector<CFResult> out(std::max(nblock, 0));
    const u64 d = H.dim();
    if (d == 0 || nblock <= 0) return out;
    hipStream_t st = H.stream;
    BlockRec rec;
    rec.setup({{&H, nblock}}, st);
    QK_HIP(hipMemcpyAsync(rec.bufU.get(), phi0_dev, sizeof(cplx) * d * nblock,
                          hipMemcpyDeviceToDevice, st));
    const std::vector<double> nrm = rec.start(1e-15);
    std::vector<double> ah, bh;
    rec.run(n_iter, ah, bh);
    for (int j = 0; j < nblock; ++j) {
        if (nrm[j] < 1e-15) {   // dynamics.cpp:26
            out[j].norm_phi0 = 0.0;
            continue;
        }
        out[j].norm_phi0 = nrm[j];
        extract(ah, bh, n_iter, nblock, j, out[j].alphas, out[j].betas);
    }
    return out;
}

std::vector<CFResult> continued_fraction_block(const Hamiltonian& H, const cplx* phi0, int nblock,
                                               int n_iter) {
    const u64 d = H.dim();
    if (d == 0 || nblock <= 0) return std::vector<CFResult>(std::max(nblock, 0));
    DevBuf buf(sizeof(cplx) * d * nblock);
    QK_HIP(hipMemcpyAsync(buf.get(), phi0, sizeof(cplx) * d * nblock, hipMemcpyHostToDevice, H.stream));
    QK_HIP(hipStreamSynchronize(H.stream));
    return continued_fraction_block_dev(H, buf.get<cplx>(), nblock, n_iter);
}

double spectral_function(const double* alphas, const double* betas, size_t n, double norm_phi0,
                         double omega, double E0, double eta) {
    // dynamics.cpp:61-89: z = omega + E0 + i eta, backward recursion
    if (n == 0) return 0.0;
    const double zr = omega + E0, zi = eta;
    double fr = 0.0, fi = 0.0;
    for (i64 i = i64(n) - 1; i >= 0; --i) {
        double dr = zr - alphas[i], di = zi;
        if (i != i64(n) - 1) {
            const double b2 = betas[i] * betas[i];
            dr -= b2 * fr;
            di -= b2 * fi;
        }
        // f = 1 / (dr + i di)
        const double den = dr * dr + di * di;
        fr = dr / den;
        fi = -di / den;
    }
    const double n2 = norm_phi0 * norm_phi0;
    return -1.0 / M_PI * (n2 * fi);
}

// ---------------------------------------------------------------------------
// FTLM (reference ftlm.cpp:122-170) over sectors and a beta grid.
// ---------------------------------------------------------------------------
std::vector<FTLMPoint> ftlm_sweep(const std::vector<const Hamiltonian*>& hs,
                                  const std::vector<double>& betas, int n_random, int n_steps,
                                  int mode) {
    const int nsec = int(hs.size()), nb = int(betas.size());
    std::vector<FTLMPoint> out(size_t(nsec) * nb);
    for (int s = 0; s < nsec; ++s)
        for (int b = 0; b < nb; ++b) out[size_t(s) * nb + b] = FTLMPoint{betas[b], 0.0, 0.0, 0.0};
    if (nsec == 0 || nb == 0) return out;
    u64 dmax = 0;
    for (auto* h : hs) dmax = std::max(dmax, h->dim());
    if (dmax == 0) return out;
    const int R = std::max(n_random, 0);

    // The reference re-seeds std::mt19937(42) on every call, so sector s uses
    // the first R*d_s complex draws of one common stream.
    Timer tm("ftlm");
    const std::shared_ptr<const cplx> stream_p = reference_normals(42, size_t(R) * dmax);
    const cplx* stream = stream_p.get();
    hipStream_t st = hs[0]->stream;
    const cplx* dstream = reference_normals_dev(42, size_t(R) * dmax, stream, st);
    const hipMemcpyKind kind = dstream ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (dstream) stream = dstream;
    tm.mark("random stream");

    // per (sector, vector) tridiagonal results and their eigensolves
    struct Sample {
        int sec;
        double nrm;
        std::vector<double> a, b, eig, z0;
    };
    std::vector<Sample> samples;
    size_t solved = 0;
    // small eigensolves of the samples not solved yet (:151-152)
    auto solve_new = [&]() {
        const size_t from = solved;
        parallel_for(int(samples.size() - from), [&](int k) {
            Sample& sm = samples[from + size_t(k)];
            const int n = int(sm.a.size());
            sm.eig.resize(n);
            sm.z0.resize(n);
            tridiag_ql(sm.a.data(), sm.b.data(), n, sm.eig.data(), sm.z0.data(), nullptr);
        });
        solved = samples.size();
    };
    for (auto* h : hs) QK_HIP(hipStreamSynchronize(h->stream));

    if (mode == 0) {
        std::vector<std::pair<const Hamiltonian*, int>> g;
        std::vector<int> secs;
        for (int s = 0; s < nsec; ++s)
            if (hs[s]->dim() > 0 && R > 0) {
                g.push_back({hs[s], R});
                secs.push_back(s);
            }
        if (!g.empty()) {
            // Two recurrences over halves of the sectors (by work, largest
            // first to the lighter half): the host solves the first half's
            // tridiagonals while the GPU runs the second half's steps.  Each
            // sample's arithmetic is its own, whichever half it is in.
            std::vector<int> half(g.size(), 0);
            if (g.size() >= 2) {
                std::vector<size_t> order(g.size());
                for (size_t i = 0; i < g.size(); ++i) order[i] = i;
                std::stable_sort(order.begin(), order.end(),
                                 [&](size_t a, size_t b) { return g[a].first->dim() > g[b].first->dim(); });
                u64 load[2] = {0, 0};
                for (size_t i : order) {
                    const int hh = load[1] < load[0] ? 1 : 0;
                    half[i] = hh;
                    load[hh] += g[i].first->dim();
                }
            }
            BlockRec rec[2];
            std::vector<size_t> members[2];
            std::vector<double> nrm[2];
            for (int hh = 0; hh < 2; ++hh) {
                std::vector<std::pair<const Hamiltonian*, int>> gh;
                for (size_t gi = 0; gi < g.size(); ++gi)
                    if (half[gi] == hh) {
                        gh.push_back(g[gi]);
                        members[hh].push_back(gi);
                    }
                if (gh.empty()) continue;
                rec[hh].setup(gh, st);
                for (size_t k = 0; k < gh.size(); ++k)
                    QK_HIP(hipMemcpyAsync(rec[hh].bufU.get<cplx>() + rec[hh].grps[k].off, stream,
                                          sizeof(cplx) * gh[k].first->dim() * R, kind, st));
                nrm[hh] = rec[hh].start(0.0);
                rec[hh].prepare(n_steps, /*graph=*/true);
            }
            tm.mark("stage vectors");
            for (int hh = 0; hh < 2; ++hh)
                if (!members[hh].empty()) rec[hh].launch();
            hipStream_t cs = nullptr;
            QK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
            try {
                for (int hh = 0; hh < 2; ++hh) {
                    if (members[hh].empty()) continue;
                    std::vector<double> ah, bh;
                    rec[hh].collect(ah, bh, cs);
                    int j = 0;
                    for (size_t gi : members[hh])
                        for (int r = 0; r < R; ++r, ++j) {
                            Sample smp{secs[gi], nrm[hh][j], {}, {}, {}, {}};
                            extract(ah, bh, n_steps, rec[hh].nsamp, j, smp.a, smp.b);
                            samples.push_back(std::move(smp));
                        }
                    solve_new();   // overlaps the other half's steps on the GPU
                }
            } catch (...) {
                (void)hipStreamSynchronize(st);
                (void)hipStreamDestroy(cs);
                throw;
            }
            QK_HIP(hipStreamDestroy(cs));
            tm.mark("block lanczos + eigensolves");
        }
    } else {
        for (int s = 0; s < nsec; ++s) {
            const u64 d = hs[s]->dim();
            if (d == 0) continue;
            BlockRec rec;
            rec.setup({{hs[s], 1}}, st, /*allow_fused=*/mode == 1 || mode == 3,
                      /*multi_single=*/mode == 1);
            for (int r = 0; r < R; ++r) {
                QK_HIP(hipMemcpyAsync(rec.bufU.get(), stream + size_t(r) * d, sizeof(cplx) * d, kind, st));
                const std::vector<double> nrm = rec.start(0.0);
                std::vector<double> ah, bh;
                rec.run(n_steps, ah, bh);
                Sample smp{s, nrm[0], {}, {}, {}, {}};
                extract(ah, bh, n_steps, 1, 0, smp.a, smp.b);
                samples.push_back(std::move(smp));
            }
        }
    }

    tm.mark("lanczos (all)");
    // small eigensolves (those not done yet), then Boltzmann weights for
    // every beta (:151-167)
    solve_new();
    tm.mark("tridiagonal eigensolves");
    std::vector<double> Zs(size_t(nsec) * nb, 0.0), Es(Zs.size(), 0.0), E2s(Zs.size(), 0.0);
    // The sums of one (sector, beta) run in sample order in registers and are
    // stored when the sector changes: threads of neighbouring betas never
    // write the same cache line inside the loop.
    parallel_for(nb, [&](int b) {
        const double beta = betas[b];
        int cur = -1;
        double Z = 0.0, E = 0.0, E2 = 0.0;
        auto flush = [&]() {
            if (cur < 0) return;
            const size_t o = size_t(cur) * nb + b;
            Zs[o] = Z;
            Es[o] = E;
            E2s[o] = E2;
        };
        for (size_t i = 0; i < samples.size(); ++i) {
            const int s = samples[i].sec;
            if (s != cur) {
                flush();
                cur = s;
                const size_t o = size_t(s) * nb + b;
                Z = Zs[o];
                E = Es[o];
                E2 = E2s[o];
            }
            const double n2 = samples[i].nrm * samples[i].nrm;
            const std::vector<double>& eig = samples[i].eig;
            const std::vector<double>& z0 = samples[i].z0;
            for (size_t k = 0; k < eig.size(); ++k) {
                const double w = n2 * z0[k] * z0[k] * std::exp(-beta * eig[k]);
                Z += w;
                E += eig[k] * w;
                E2 += eig[k] * eig[k] * w;
            }
        }
        flush();
    });
    for (int s = 0; s < nsec; ++s) {
        if (hs[s]->dim() == 0) continue;
        for (int b = 0; b < nb; ++b) {
            const size_t o = size_t(s) * nb + b;
            const double beta = betas[b];
            const double Z = Zs[o] / R, E = Es[o] / R, E2 = E2s[o] / R;
            out[o] = FTLMPoint{beta, Z, E / Z, (beta * beta) * (E2 / Z - (E / Z) * (E / Z))};
        }
    }
    return out;
}

}  // namespace qk
