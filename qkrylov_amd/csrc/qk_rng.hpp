// Reference random streams (see qk_rng.cpp).
#pragma once

#include "qk_common.hpp"

#include <cstdint>
#include <random>
#include <vector>

namespace qk {

// MT19937 (Matsumoto and Nishimura, ACM TOMACS 8(1):3-30, 1998) as a
// uniform random bit generator: word for word the sequence of std::mt19937
// for the same seed, produced 624 words at a time by loops the compiler
// vectorises (the standard engine twists one word per call).  The
// distributions below read it exactly as they would read std::mt19937.
class BlockMT19937 {
public:
    using result_type = std::uint32_t;
    static constexpr result_type min() { return 0u; }
    static constexpr result_type max() { return 0xffffffffu; }
    explicit BlockMT19937(result_type seed);
    result_type operator()() {
        if (pos_ == kN) refill();
        return out_[pos_++];
    }
    // the next n words, as n calls would return them
    void fill(std::uint32_t* w, std::size_t n);

private:
    static constexpr int kN = 624;
    void refill();
    std::uint32_t s_[kN];
    std::uint32_t out_[kN];
    int pos_ = kN;
};

// std::mt19937(seed), uniform(-1, 1), per element imaginary then real,
// written to out[0..n) (host memory, no initialisation needed).
void uniform_start(u64 n, unsigned seed, cplx* out);

// std::mt19937(seed) + std::normal_distribution<double>(0, 1), one stream
// consumed element by element (imaginary then real), as ftlm() draws its
// n_random vectors one after another.
class NormalStream {
public:
    explicit NormalStream(unsigned seed);
    void fill(cplx* out, u64 n);

private:
    BlockMT19937 rng_;
    std::normal_distribution<double> dist_;
};

}  // namespace qk
