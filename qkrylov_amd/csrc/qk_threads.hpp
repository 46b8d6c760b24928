// Host worker threads for the library's CPU-side loops (random-stream
// conversion, tridiagonal eigensolves, training-pair staging).
//
// run_ranges(n, f) splits [0, n) into contiguous ranges, one per thread
// (at most 16, at most the hardware threads), and calls f(begin, end) on
// each.  If the system refuses a thread (std::thread's constructor throws,
// e.g. EAGAIN under a process limit), the threads already started are
// joined and the ranges that have no thread run on the calling thread, so
// the result is the same and no joinable std::thread is ever destroyed
// (which would call std::terminate).  An exception thrown by f on a worker
// is rethrown on the caller after every worker has been joined.
#pragma once

#include <algorithm>
#include <cstdint>
#include <exception>
#include <mutex>
#include <system_error>
#include <thread>
#include <vector>

namespace qk {

inline unsigned host_threads(std::uint64_t n_items) {
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned nt = std::max(1u, std::min(hw ? hw : 1u, 16u));
    return unsigned(std::max<std::uint64_t>(1, std::min<std::uint64_t>(nt, n_items)));
}

template <typename F>
void run_ranges(std::uint64_t n, unsigned nt, F&& f) {
    nt = unsigned(std::max<std::uint64_t>(1, std::min<std::uint64_t>(nt, n)));
    if (nt <= 1) {
        if (n) f(std::uint64_t(0), n);
        return;
    }
    std::exception_ptr err;
    std::mutex err_mu;
    auto body = [&](std::uint64_t a, std::uint64_t b) {
        try {
            f(a, b);
        } catch (...) {
            std::lock_guard<std::mutex> g(err_mu);
            if (!err) err = std::current_exception();
        }
    };
    std::vector<std::thread> th;
    th.reserve(nt);
    unsigned started = 0;
    try {
        for (; started < nt; ++started)
            th.emplace_back(body, n * started / nt, n * (started + 1) / nt);
    } catch (...) {
        // fewer threads than asked for: the rest runs here
    }
    for (unsigned t = started; t < nt; ++t) body(n * t / nt, n * (t + 1) / nt);
    for (auto& t : th) t.join();
    if (err) std::rethrow_exception(err);
}

}  // namespace qk
