// Host random streams of the reference solvers: libstdc++'s distributions,
// called exactly as the reference calls them, over an MT19937 engine that
// produces std::mt19937's words (BlockMT19937, a block at a time).
// The reference writes KComplex(dist(rng), dist(rng)); its GCC build
// evaluates constructor arguments right to left, so for every element the
// imaginary part is drawn first and the real part second.  Compiled with
// -ffp-contract=off so the polar-method arithmetic inside
// std::normal_distribution rounds as on the reference's baseline build.
#include "qk_rng.hpp"
#include "qk_threads.hpp"

#include <random>
#include <cstring>
#include <algorithm>
#include <cmath>

namespace qk {

BlockMT19937::BlockMT19937(result_type seed) {
    // the standard initialisation (std::mt19937's seed(value))
    s_[0] = seed;
    for (int i = 1; i < kN; ++i) s_[i] = 1812433253u * (s_[i - 1] ^ (s_[i - 1] >> 30)) + std::uint32_t(i);
}

// The twist and the conversion loops are built twice, for AVX2 and for the
// baseline x86-64 ISA, and the loader picks the clone the host supports
// (GNU ifunc); both clones round identically (IEEE adds, multiplies and
// divides, no contraction), the AVX2 one only vectorises wider.
#define QK_VECTOR_CLONES __attribute__((target_clones("avx2", "default")))

namespace {

// one block: renew the 624 state words and temper them into out
QK_VECTOR_CLONES void twist_block(std::uint32_t* s, std::uint32_t* out) {
    constexpr int kN = 624;
    // twist: word i takes the top bit of word i and the low 31 bits of word
    // i+1, and word i+397 (mod 624) -- the old one for i < 227, the one
    // already renewed in this block after that; then temper
    auto tw = [](std::uint32_t a, std::uint32_t b, std::uint32_t c) {
        const std::uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
        return c ^ (y >> 1) ^ ((0u - (y & 1u)) & 0x9908b0dfu);
    };
    int i = 0;
    for (; i < kN - 397; ++i) s[i] = tw(s[i], s[i + 1], s[i + 397]);
    for (; i < kN - 1; ++i) s[i] = tw(s[i], s[i + 1], s[i + 397 - kN]);
    s[kN - 1] = tw(s[kN - 1], s[0], s[396]);
    for (int k = 0; k < kN; ++k) {
        std::uint32_t y = s[k];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        out[k] = y;
    }
}

}  // namespace

void BlockMT19937::refill() {
    twist_block(s_, out_);
    pos_ = 0;
}

void BlockMT19937::fill(std::uint32_t* w, std::size_t n) {
    std::size_t k = 0;
    while (k < n) {
        if (pos_ == kN) refill();
        const std::size_t take = std::min<std::size_t>(n - k, std::size_t(kN - pos_));
        std::memcpy(w + k, out_ + pos_, take * sizeof(std::uint32_t));
        pos_ += int(take);
        k += take;
    }
}

namespace {

// std::uniform_real_distribution<double>(-1, 1) over a 32-bit engine:
// generate_canonical<double, 53> takes two words, S = w0 + w1 * 2^32
// (one rounding), u = S / 2^64 (below 1 by construction, else the largest
// double below 1), and the distribution returns u * (b - a) + a
// ([rand.util.canonical], [rand.dist.uni.real]; no contraction here).
inline double uniform_pm1(std::uint32_t w0, std::uint32_t w1) {
    double sum = double(w0);
    sum += double(w1) * 4294967296.0;
    double u = sum / 18446744073709551616.0;
    if (u >= 1.0) u = std::nextafter(1.0, 0.0);
    return u * 2.0 + -1.0;
}

// elements [a, b) of a chunk from its words, element i from words 4i..4i+3
QK_VECTOR_CLONES void convert_uniform(const std::uint32_t* w, cplx* v, u64 a, u64 b) {
    for (u64 i = a; i < b; ++i) {
        const double im = uniform_pm1(w[4 * i], w[4 * i + 1]);
        const double re = uniform_pm1(w[4 * i + 2], w[4 * i + 3]);
        v[i] = cplx{re, im};
    }
}

}  // namespace

void uniform_start(u64 n, unsigned seed, cplx* v) {
    // lanczos.cpp:115-120 (also davidson.cpp:106-107, 128).  The words come
    // from one engine in order; element i takes words 4i..4i+3 (imaginary
    // part first), so once a chunk's words are drawn its elements are
    // independent and are converted on all host threads.
    BlockMT19937 rng(seed);
    const u64 chunk = u64(1) << 20;
    std::vector<std::uint32_t> w(4 * std::min<u64>(n, chunk));
    const unsigned nth = host_threads(chunk);
    for (u64 c0 = 0; c0 < n; c0 += chunk) {
        const u64 len = std::min(chunk, n - c0);
        rng.fill(w.data(), 4 * len);
        run_ranges(len, len < 65536 ? 1u : nth,
                   [&](u64 a, u64 b) { convert_uniform(w.data(), v + c0, a, b); });
    }
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
