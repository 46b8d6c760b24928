// Host random streams of the reference solvers, generated with libstdc++'s
// std::mt19937 and distributions exactly as the reference calls them.
// The reference writes KComplex(dist(rng), dist(rng)); its GCC build
// evaluates constructor arguments right to left, so for every element the
// imaginary part is drawn first and the real part second.  Compiled with
// -ffp-contract=off so the polar-method arithmetic inside
// std::normal_distribution rounds as on the reference's baseline build.
#include "qk_rng.hpp"

#include <random>

namespace qk {

std::vector<cplx> uniform_start(u64 n, unsigned seed) {
    // lanczos.cpp:115-120 (also davidson.cpp:106-107, 128)
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> dist(-1.0, 1.0);
    std::vector<cplx> v(n);
    for (u64 i = 0; i < n; ++i) {
        const double im = dist(rng);
        const double re = dist(rng);
        v[i] = cplx{re, im};
    }
    return v;
}

NormalStream::NormalStream(unsigned seed) : rng_(seed), dist_(0.0, 1.0) {}

void NormalStream::fill(cplx* out, u64 n) {
    // ftlm.cpp:133-143
    for (u64 i = 0; i < n; ++i) {
        const double im = dist_(rng_);
        const double re = dist_(rng_);
        out[i] = cplx{re, im};
    }
}

}  // namespace qk
