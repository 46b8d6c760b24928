// Common types for the qkrylov MI355X library: complex FP64, state words,
// error reporting.  No torch, no CUDA, no Kokkos.
#pragma once

#include <cstdint>
#include <cstddef>
#include <stdexcept>
#include <string>

namespace qk {

using u64 = std::uint64_t;
using u32 = std::uint32_t;
using u16 = std::uint16_t;
using i64 = std::int64_t;

// Interleaved complex double, layout-compatible with std::complex<double>
// and with the reference's Kokkos::complex<double> (16 B, re then im).
struct cplx {
    double re, im;
};

// Complex product written out term by term: (a.re*b.re - a.im*b.im,
// a.re*b.im + a.im*b.re).  For finite operands this is what the reference's
// std::complex<double>::operator*= evaluates (no FMA on the reference's
// baseline x86-64 build), so compiling host code that uses it with
// -ffp-contract=off reproduces the reference's matrix elements bit for bit.
inline cplx cmul(cplx a, cplx b) {
    return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
inline cplx cadd(cplx a, cplx b) { return cplx{a.re + b.re, a.im + b.im}; }
inline cplx cconj(cplx a) { return cplx{a.re, -a.im}; }

inline int popc(u64 x) { return __builtin_popcountll(x); }

// Every failure inside the library is a qk::Error; the C ABI converts it to
// QKRYLOV_ERROR_EXCEPTION (or a NULL handle) and keeps the message for
// qkrylov_last_error().
struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

[[noreturn]] void fail(const std::string& msg);

}  // namespace qk
