"""Pins the draw order of the reference's random start vectors on the CPU.

The reference fills its start vectors with `KComplex(dist(rng), dist(rng))`
(src/solvers/lanczos.cpp:114-118, davidson.cpp:106-107 and :128,
ftlm.cpp:133-143): two calls of one distribution as the two arguments of a
complex constructor, whose evaluation order C++ leaves to the compiler.  The
program below uses that expression shape, compiled with the host g++ (the
compiler of the reference's CI builds), for two stand-ins of the complex
type: a plain two-argument constructor and one templated on its argument
types and taking them by const reference (the form Kokkos::complex
declares).  Its streams must equal, bit for bit, the library's
qkrylov_reference_{start,normals}_f64 and the oracle's streams, which both
draw the imaginary part of an element first.
"""
import shutil
import subprocess

import numpy as np
import pytest

PROGRAM = r"""
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

struct PlainComplex {
    double re, im;
    PlainComplex() : re(0), im(0) {}
    PlainComplex(double r, double i) : re(r), im(i) {}
};

struct TemplComplex {
    double re, im;
    TemplComplex() : re(0), im(0) {}
    template <class R1, class R2>
    TemplComplex(const R1& r, const R2& i) : re(r), im(i) {}
};

template <class C, class Dist>
void fill(std::vector<C>& v, unsigned seed, Dist dist) {
    std::mt19937 rng(seed);
    for (size_t i = 0; i < v.size(); ++i) v[i] = C(dist(rng), dist(rng));
}

int main(int argc, char** argv) {
    const size_t n = std::strtoull(argv[1], nullptr, 10);
    const unsigned seed = unsigned(std::strtoul(argv[2], nullptr, 10));
    const int normal = std::atoi(argv[3]);
    std::vector<PlainComplex> a(n);
    std::vector<TemplComplex> b(n);
    if (normal) {
        fill(a, seed, std::normal_distribution<double>(0.0, 1.0));
        fill(b, seed, std::normal_distribution<double>(0.0, 1.0));
    } else {
        fill(a, seed, std::uniform_real_distribution<double>(-1.0, 1.0));
        fill(b, seed, std::uniform_real_distribution<double>(-1.0, 1.0));
    }
    std::fwrite(a.data(), sizeof(PlainComplex), n, stdout);
    std::fwrite(b.data(), sizeof(TemplComplex), n, stdout);
    return 0;
}
"""


@pytest.fixture(scope="module")
def stream_prog(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host g++")
    d = tmp_path_factory.mktemp("rngorder")
    src = d / "order.cpp"
    src.write_text(PROGRAM)
    exe = d / "order"
    subprocess.run([gxx, "-O2", "-std=c++20", "-o", str(exe), str(src)], check=True)
    return exe


def _draw(exe, n, seed, normal):
    out = subprocess.run([str(exe), str(n), str(seed), str(int(normal))], check=True,
                         capture_output=True).stdout
    v = np.frombuffer(out, dtype=np.float64).reshape(2, n, 2)
    return v[0], v[1]          # (re, im) pairs of the two stand-ins


@pytest.mark.parametrize("normal,seed", [(False, 1234), (False, 7), (True, 42), (True, 3)])
def test_expression_shape_matches_streams(stream_prog, qk, orc, normal, seed):
    n = 4099
    plain, templ = _draw(stream_prog, n, seed, normal)
    assert np.array_equal(plain, templ)
    lib = qk.load()
    got = np.empty(n, np.complex128)
    if normal:
        assert lib.qkrylov_reference_normals_f64(n, seed, got.ctypes.data) == 0
        oracle = orc.normal_stream(n, seed)
    else:
        assert lib.qkrylov_reference_start_f64(n, seed, got.ctypes.data) == 0
        oracle = orc.uniform_start(n, seed)
    assert np.array_equal(got.view(np.float64).reshape(n, 2), plain)
    assert np.array_equal(oracle.view(np.float64).reshape(n, 2), plain)


def test_imaginary_part_is_the_first_draw(stream_prog):
    """Independent of both implementations: element 0's imaginary part is
    the first number of std::mt19937(1234) through uniform(-1, 1)."""
    plain, _ = _draw(stream_prog, 2, 1234, False)
    # std::mt19937(1234) first outputs, through generate_canonical<double, 53>
    # (two 32-bit words w0, w1: (w0 + w1 * 2^32) / 2^64), then 2u - 1
    mt = np.random.RandomState(1234)       # MT19937 with the same init_genrand seeding
    w = mt.randint(0, 2**32, size=4, dtype=np.uint64).astype(np.float64)
    u0 = (w[0] + w[1] * 2.0**32) / 2.0**64
    u1 = (w[2] + w[3] * 2.0**32) / 2.0**64
    assert plain[0, 1] == u0 * 2.0 - 1.0
    assert plain[0, 0] == u1 * 2.0 - 1.0
