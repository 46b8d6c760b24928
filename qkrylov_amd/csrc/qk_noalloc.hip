// qkrylov_block_dot_noalloc_f64: the FP64-MFMA projection G = V^H W of
// qkrylov_block_dot_f64 for callers that cannot allocate device memory
// (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc).  Inputs and output are the caller's device buffers; the
// kernels' scratch -- the tables of vector pointers and the per-workgroup
// partial sums -- lives in one file-scope device array, `scratch`, of
// kScratchBytes per GPU, held until the process ends.  No hipMalloc,
// hipFree or allocating library call is made.
//
// The array's name in host code is only a stand-in for the symbol: its
// device address is looked up once per device (hipGetSymbolAddress, with
// that device current) and only that address is handed to kernels.  Every
// call on one GPU shares that GPU's array: calls that overlap in time on one
// GPU race (they must be ordered, e.g. on one stream).  The kernels write
// every scratch byte they read in the same call, so the previous call's
// contents never leak into a result.
#include "qk_vec.hpp"

#include "qkrylov/c_api.h"

#include <mutex>

namespace qk {

// from c_api.cpp: the calling thread's message for qkrylov_last_error()
void capi_set_error(const char* msg);

namespace {

constexpr size_t kPtrBytes = 1024;   // 64 V pointers + 16 W pointers, rounded up
// partial sums: at most 1024 workgroups x (4 x 16 x 16) complex values
constexpr size_t kPartBytes = size_t(1024) * 1024 * sizeof(cplx);
constexpr size_t kScratchBytes = kPtrBytes + kPartBytes;   // 16 MiB + 1 KiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// this device's address of `scratch` (looked up on the first call per device)
unsigned char* scratch_ptr() {
    int dev = 0;
    QK_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) fail("block_dot_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            fail(std::string("block_dot_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
        g_scratch[dev] = p;
    }
    return static_cast<unsigned char*>(g_scratch[dev]);
}

// V_i = V + i*ldv, W_j = W + j*ldw into the pointer tables
__global__ void k_noalloc_ptrs(const cplx** tab, const cplx* V, u64 ldv, int m, const cplx* W, u64 ldw,
                               int nw) {
    const int i = int(threadIdx.x);
    if (i < m) tab[i] = V + u64(i) * ldv;
    if (i < nw) tab[64 + i] = W + u64(i) * ldw;
}

}  // namespace
}  // namespace qk

extern "C" int qkrylov_block_dot_noalloc_f64(const void* V_dev, uint64_t ldv, int m, const void* W_dev,
                                             uint64_t ldw, int nw, uint64_t n, void* G_dev,
                                             void* stream) {
    using namespace qk;
    if (!V_dev || !W_dev || !G_dev || m <= 0 || nw <= 0 || m > 64 || nw > 16 || ldv < n || ldw < n) {
        capi_set_error("block_dot_noalloc: need V, W, G, 1 <= m <= 64, 1 <= nw <= 16, ldv >= n, ldw >= n");
        return QKRYLOV_ERROR_INVALID_ARG;
    }
    const size_t need = vec::block_dot_scratch_bytes(m, n);
    if (need > kPartBytes) {
        capi_set_error("block_dot_noalloc: the partial sums exceed the 16 MiB scratch array");
        return QKRYLOV_ERROR_INVALID_ARG;
    }
    try {
        unsigned char* base = scratch_ptr();
        const cplx** tab = reinterpret_cast<const cplx**>(base);
        cplx* part = reinterpret_cast<cplx*>(base + kPtrBytes);
        hipStream_t st = static_cast<hipStream_t>(stream);
        hipLaunchKernelGGL(k_noalloc_ptrs, dim3(1), dim3(64), 0, st, tab, static_cast<const cplx*>(V_dev),
                           u64(ldv), m, static_cast<const cplx*>(W_dev), u64(ldw), nw);
        QK_HIP(hipGetLastError());
        vec::block_dot_into(tab, m, tab + 64, nw, n, static_cast<cplx*>(G_dev), part, st);
        return QKRYLOV_SUCCESS;
    } catch (const std::exception& e) {
        capi_set_error(e.what());
        return QKRYLOV_ERROR_EXCEPTION;
    } catch (...) {
        capi_set_error("unknown exception");
        return QKRYLOV_ERROR_EXCEPTION;
    }
}
