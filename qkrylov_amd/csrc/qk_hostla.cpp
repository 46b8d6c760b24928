// Host small eigensolvers (see qk_hostla.hpp).  Compiled with
// -ffp-contract=off: every product and sum is rounded separately, as on the
// reference's baseline x86-64 build, so the tridiagonal eigenvalues match the
// reference's QL sweep operation for operation.
#include "qk_hostla.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <numeric>

namespace qk {

void tridiag_ql(const double* a, const double* b, int n, double* eig, double* z0, double* Z) {
    if (n <= 0) return;
    if (n == 1) {
        eig[0] = a[0];
        if (z0) z0[0] = 1.0;
        if (Z) Z[0] = 1.0;
        return;
    }
    std::vector<double> d(a, a + n), e(n, 0.0);
    for (int i = 0; i + 1 < n; ++i) e[i] = b[i];
    std::vector<double> row0;
    if (z0) {
        row0.assign(n, 0.0);
        row0[0] = 1.0;
    }
    if (Z) {
        std::fill(Z, Z + size_t(n) * n, 0.0);
        for (int i = 0; i < n; ++i) Z[size_t(i) * n + i] = 1.0;
    }

    for (int sweep = 0; sweep < 1000; ++sweep) {
        // deflate negligible couplings
        for (int i = 0; i + 1 < n; ++i)
            if (std::abs(e[i]) < 1e-14 * (std::abs(d[i]) + std::abs(d[i + 1]))) e[i] = 0.0;
        // trailing unreduced block [lo, hi]
        int hi = n - 1;
        while (hi > 0 && e[hi - 1] == 0.0) --hi;
        if (hi == 0) break;
        int lo = hi - 1;
        while (lo > 0 && e[lo - 1] != 0.0) --lo;

        // shift from the trailing 2x2 (eigenvalue nearer d[hi])
        const double half = (d[hi - 1] - d[hi]) / 2.0;
        const double cc = e[hi - 1] * e[hi - 1];
        const double rt = std::sqrt(half * half + cc);
        const double mu = (half > 0) ? d[hi] - cc / (half + rt) : d[hi] - cc / (half - rt);

        double x = d[lo] - mu;
        double y = e[lo];
        for (int i = lo; i < hi; ++i) {
            const double r = std::hypot(x, y);
            const double cs = x / r;
            const double sn = y / r;
            if (i > lo) e[i - 1] = r;
            const double f = cs * d[i] + sn * e[i];
            const double g = cs * e[i] + sn * d[i + 1];
            const double h = sn * d[i] - cs * e[i];
            const double k = sn * e[i] - cs * d[i + 1];
            d[i] = cs * f + sn * g;
            e[i] = cs * h + sn * k;
            d[i + 1] = sn * h - cs * k;
            if (z0) {
                const double z1 = row0[i], z2 = row0[i + 1];
                row0[i] = cs * z1 + sn * z2;
                row0[i + 1] = sn * z1 - cs * z2;
            }
            if (Z) {
                for (int j = 0; j < n; ++j) {
                    double* zr = Z + size_t(j) * n;
                    const double z1 = zr[i], z2 = zr[i + 1];
                    zr[i] = cs * z1 + sn * z2;
                    zr[i + 1] = sn * z1 - cs * z2;
                }
            }
            if (i < hi - 1) {
                x = e[i];
                y = sn * e[i + 1];
                e[i + 1] = -cs * e[i + 1];
            }
        }
    }
    for (int i = 0; i < n; ++i) eig[i] = d[i];
    if (z0)
        for (int i = 0; i < n; ++i) z0[i] = row0[i];
}

// Householder reduction to tridiagonal form (A = Q T Q^H, T Hermitian
// tridiagonal), a diagonal unitary D that makes T's off-diagonal real and
// non-negative (T = D T' D^H), and tridiag_ql on T'; the eigenvectors are
// Q D z.  About 4/3 n^3 complex multiply-adds for the reduction plus the QL
// sweeps, against the several full sweeps of rotations Jacobi needs.
void hermitian_eigh(const std::vector<cplx>& Ain, int n, std::vector<double>& evals,
                    std::vector<cplx>& vecs) {
    using C = std::complex<double>;
    evals.assign(n, 0.0);
    vecs.assign(size_t(n) * n, cplx{0.0, 0.0});
    if (n <= 0) return;
    std::vector<C> A(size_t(n) * n), Q(size_t(n) * n, C(0.0, 0.0)), v(n), w(n);
    for (size_t i = 0; i < A.size(); ++i) A[i] = C(Ain[i].re, Ain[i].im);
    for (int i = 0; i < n; ++i) Q[size_t(i) * n + i] = 1.0;
    auto a = [&](int i, int j) -> C& { return A[size_t(i) * n + j]; };
    for (int k = 0; k + 2 < n; ++k) {
        // reflect x = A[k+1.., k] onto alpha e_{k+1}, alpha = -phase(x0) |x|
        double xn2 = 0.0;
        for (int i = k + 1; i < n; ++i) xn2 += std::norm(a(i, k));
        const double xn = std::sqrt(xn2);
        if (xn == 0.0) continue;
        const C x0 = a(k + 1, k);
        const double ax0 = std::abs(x0);
        const C phase = ax0 > 0.0 ? x0 / ax0 : C(1.0, 0.0);
        std::fill(v.begin(), v.end(), C(0.0, 0.0));
        v[k + 1] = x0 + phase * xn;
        for (int i = k + 2; i < n; ++i) v[i] = a(i, k);
        double vn2 = 0.0;
        for (int i = k + 1; i < n; ++i) vn2 += std::norm(v[i]);
        if (vn2 == 0.0) continue;
        const double tau = 2.0 / vn2;
        // A <- H A H, H = I - tau v v^H:  p = tau A v, K = tau/2 v^H p,
        // q = p - K v, A <- A - v q^H - q v^H (rows and columns k..n-1)
        for (int i = k; i < n; ++i) {
            C s(0.0, 0.0);
            for (int j = k + 1; j < n; ++j) s += a(i, j) * v[j];
            w[i] = tau * s;
        }
        C K(0.0, 0.0);
        for (int i = k + 1; i < n; ++i) K += std::conj(v[i]) * w[i];
        K *= 0.5 * tau;
        for (int i = k; i < n; ++i) w[i] -= K * v[i];
        for (int i = k; i < n; ++i)
            for (int j = k; j < n; ++j) a(i, j) -= v[i] * std::conj(w[j]) + w[i] * std::conj(v[j]);
        // Q <- Q H
        for (int i = 0; i < n; ++i) {
            C s(0.0, 0.0);
            for (int j = k + 1; j < n; ++j) s += Q[size_t(i) * n + j] * v[j];
            s *= tau;
            for (int j = k + 1; j < n; ++j) Q[size_t(i) * n + j] -= s * std::conj(v[j]);
        }
    }
    // T' = D^H T D with phi_0 = 1, phi_{i+1} = phi_i t_i / |t_i|
    std::vector<double> d(n), e(std::max(n - 1, 1), 0.0);
    std::vector<C> phi(n, C(1.0, 0.0));
    for (int i = 0; i < n; ++i) d[i] = a(i, i).real();
    for (int i = 0; i + 1 < n; ++i) {
        const C t = a(i + 1, i);
        const double at = std::abs(t);
        e[i] = at;
        phi[i + 1] = at > 0.0 ? phi[i] * t / at : phi[i];
    }
    std::vector<double> eig(n), Z(size_t(n) * n);
    tridiag_ql(d.data(), e.data(), n, eig.data(), nullptr, Z.data());
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return eig[x] < eig[y]; });
    for (int k = 0; k < n; ++k) {
        const int c = ord[k];
        evals[k] = eig[c];
        for (int j = 0; j < n; ++j) {
            C s(0.0, 0.0);
            for (int l = 0; l < n; ++l) s += Q[size_t(j) * n + l] * phi[l] * Z[size_t(l) * n + c];
            vecs[size_t(k) * n + j] = cplx{s.real(), s.imag()};
        }
    }
}

double tridiag_ground(const std::vector<double>& a, const std::vector<double>& b, int n,
                      std::vector<double>* vec) {
    if (n == 0) {
        if (vec) vec->clear();
        return 0.0;
    }
    std::vector<double> eig(n);
    std::vector<double> Z;
    if (vec) Z.assign(size_t(n) * n, 0.0);
    tridiag_ql(a.data(), b.data(), n, eig.data(), nullptr, vec ? Z.data() : nullptr);
    int imin = 0;
    for (int i = 1; i < n; ++i)
        if (eig[i] < eig[imin]) imin = i;
    if (vec) {
        vec->resize(n);
        for (int j = 0; j < n; ++j) (*vec)[j] = Z[size_t(j) * n + imin];
    }
    return eig[imin];
}

}  // namespace qk
