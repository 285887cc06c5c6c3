// probe.hip — HBM streaming-read probe: the achievable read bandwidth that
// the bandwidth-bound kernels (the skinny GEMMs) are compared against.
// Not part of any GEMM: a diagnostic with no counterpart in the reference.
#include "api_common.h"

namespace blasdsp {
namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Each thread streams 16-byte vectors, 4 in flight per iteration, and XORs
// them so the loads are not removed; one word per workgroup is written.
__global__ void __launch_bounds__(256) read_probe_kernel(const u32x4 *__restrict__ p, int64_t nvec,
                                                         uint32_t *__restrict__ out)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    u32x4 acc = {0, 0, 0, 0};
    for (; i + 3 * stride < nvec; i += 4 * stride) {
        const u32x4 a = __builtin_nontemporal_load(p + i);
        const u32x4 b = __builtin_nontemporal_load(p + i + stride);
        const u32x4 c = __builtin_nontemporal_load(p + i + 2 * stride);
        const u32x4 d = __builtin_nontemporal_load(p + i + 3 * stride);
        acc ^= a ^ b ^ c ^ d;
    }
    for (; i < nvec; i += stride) acc ^= p[i];
    __shared__ uint32_t red[256];
    red[threadIdx.x] = acc[0] ^ acc[1] ^ acc[2] ^ acc[3];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] ^= red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

}  // namespace
}  // namespace blasdsp

extern "C" int blasdsp_probe_read(const void *p, int64_t bytes, void *out, int blocks, void *stream)
{
    if (!p || !out || bytes < 16 || bytes % 16 != 0 || blocks <= 0 || ((uintptr_t)p % 16) != 0)
        return BLASDSP_ERR_INVALID;
    hipLaunchKernelGGL(blasdsp::read_probe_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       (const blasdsp::u32x4 *)p, bytes / 16, (uint32_t *)out);
    return blasdsp::hip_status(hipGetLastError());
}
