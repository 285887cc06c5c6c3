// noalloc.hip — blasdsp_gemm_noalloc: the GEMM entry point that allocates
// nothing, for callers that cannot call hipMalloc (code under hipGraph
// capture, tight serving loops, harnesses that forbid allocation).
//
// No counterpart in the reference: its kernels allocate on-chip scratch per
// call and free it before returning (dgemm_release_V1.1/device_code/
// dgemm_kernel.c:2235-2241); HBM is always the caller's.
//
// It runs the same kernels as blasdsp_dgemm_ex / blasdsp_sgemm_ex (same
// choice of kernel, tile and K slices for the same arguments, hence the same
// bits).  The only scratch those kernels use is the split-K partial sums
// ([slice][M][N]); here they live in `scratch`, a file-scope device array of
// kNoallocScratchBytes that every GPU the library runs on holds until the
// process ends.  Each device's copy sits at its own address, looked up once
// per device with hipGetSymbolAddress (with that device current) and kept in
// g_scratch; the array's name is never used as an address on the host.
//
// Every slice writes every element of its partial before the reduction reads
// it, so nothing has to be cleared between calls.  All calls on one GPU share
// its array: calls that overlap on one GPU (two streams, or two host threads)
// race on it.
#include <mutex>

#include "api_common.h"

namespace blasdsp {
namespace {

constexpr size_t kNoallocScratchBytes = (size_t)BLASDSP_NOALLOC_SCRATCH_BYTES;
static_assert(kNoallocScratchBytes % 256 == 0, "scratch size is a multiple of its alignment");

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch[kMaxDevices] = {};

// The current device's copy of `scratch` (looked up on first use).
hipError_t scratch_ptr(void **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

template <typename T>
int noalloc_entry(const blasdsp_gemm_desc &d, int algo, int splitk, hipStream_t s)
{
    Gemm<T> g;
    int st = check<T>(g, d.transA, d.transB, d.M, d.N, d.K, (T)d.alpha, (const T *)d.A, d.lda,
                      (const T *)d.B, d.ldb, (T)d.beta, (T *)d.C, d.ldc);
    if (st != BLASDSP_SUCCESS) return st;
    if (g.M == 0 || g.N == 0) return BLASDSP_SUCCESS;
    if (g.K == 0 || g.alpha == T(0)) return hip_status(launch_scale(g, s));
    KernelChoice c;
    st = choose_kernel(g, algo, splitk, c);
    if (st != BLASDSP_SUCCESS) return st;
    if (kernel_scratch_bytes(g, c.splitk) > kNoallocScratchBytes) {
        // An explicit slice count whose partials do not fit is refused; a
        // count the library chose is lowered until they do (1 needs none).
        if (splitk > 0) return BLASDSP_ERR_NOMEM;
        const size_t per_slice = (size_t)g.M * (size_t)g.N * sizeof(T);
        const size_t fit = kNoallocScratchBytes / per_slice;
        c.splitk = fit >= 2 ? (int)(fit < (size_t)c.splitk ? fit : (size_t)c.splitk) : 1;
    }
    T *work = nullptr;
    if (kernel_scratch_bytes(g, c.splitk) > 0) {
        void *p = nullptr;
        const hipError_t e = scratch_ptr(&p);
        if (e != hipSuccess) return hip_status(e);
        work = (T *)p;
    }
    return launch_kernel(g, c, work, s);
}

}  // namespace
}  // namespace blasdsp

extern "C" int blasdsp_gemm_noalloc(const blasdsp_gemm_desc *desc, int algo, int splitk, void *stream)
{
    if (!desc) return BLASDSP_ERR_INVALID;
    if (desc->dtype == BLASDSP_FLOAT64)
        return blasdsp::noalloc_entry<double>(*desc, algo, splitk, (hipStream_t)stream);
    if (desc->dtype == BLASDSP_FLOAT32)
        return blasdsp::noalloc_entry<float>(*desc, algo, splitk, (hipStream_t)stream);
    return BLASDSP_ERR_INVALID;
}
