// Reference random streams (see qk_rng.cpp).
#pragma once

#include "qk_common.hpp"

#include <random>
#include <vector>

namespace qk {

// std::mt19937(seed), uniform(-1, 1), per element imaginary then real.
std::vector<cplx> uniform_start(u64 n, unsigned seed);

// std::mt19937(seed) + std::normal_distribution<double>(0, 1), one stream
// consumed element by element (imaginary then real), as ftlm() draws its
// n_random vectors one after another.
class NormalStream {
public:
    explicit NormalStream(unsigned seed);
    void fill(cplx* out, u64 n);

private:
    std::mt19937 rng_;
    std::normal_distribution<double> dist_;
};

}  // namespace qk
