// probe.hip — diagnostics with no counterpart in the reference, part of no
// GEMM:
//   read_probe_kernel   the achievable HBM read bandwidth that the
//                       bandwidth-bound kernels (the skinny GEMMs) are
//                       compared against;
//   mfma_probe_kernel   the sustained rate of each matrix-core instruction
//                       form the GEMMs use, from registers alone;
//   lds_probe_kernel    the throughput of global -> LDS operand loads from an
//                       L2-resident region, per load path.
// This file is built with the MFMAs in their VGPR form (see the Makefile), so
// the MFMA streams carry no accumulator copies.
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

typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef double d4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

constexpr int kProbeVec = 1024;  // 16-byte operand words read by the MFMA probe

// One wave per SIMD (4 per workgroup): each wave loads its operands once from
// `src` (kProbeVec 16-byte words), then issues `iters` rounds of independent
// MFMAs of one form on them (32 accumulator tiles for the 16 x 16 forms, 8 for
// the 32 x 32 forms, 16 for FP64 and FP8 16 x 16), and writes one sum per
// thread.  Forms: 0 bf16 16x16x32, 1 bf16 32x32x16, 2 f16 16x16x32, 3 f32
// 16x16x4, 4 f32 32x32x2, 5 f64 16x16x4, 6 fp8 16x16x128 and 7 fp8 32x32x64,
// both with unit scales.
template <int FORM>
__global__ void __launch_bounds__(256) mfma_probe_kernel(const u32x4 *__restrict__ src, float *__restrict__ out,
                                                         int iters)
{
    const int lane = threadIdx.x & 63;
    auto word = [&](int i) { return src[(lane * 12 + i) & (kProbeVec - 1)]; };
    float s = 0.0f;
    if constexpr (FORM == 0 || FORM == 2) {
        u32x4 a[8], b[4];
#pragma unroll
        for (int i = 0; i < 8; ++i) a[i] = word(i);
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = word(8 + j);
        f4 acc[8][4] = {};
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if constexpr (FORM == 0)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(bf16x8, a[i]), __builtin_bit_cast(bf16x8, b[j]), acc[i][j], 0, 0, 0);
                    else
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                            __builtin_bit_cast(f16x8, a[i]), __builtin_bit_cast(f16x8, b[j]), acc[i][j], 0, 0, 0);
                }
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) s += acc[i][j][0] + acc[i][j][3];
    } else if constexpr (FORM == 1) {
        u32x4 a[4], b[2];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = word(i);
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = word(8 + j);
        f16v acc[4][2] = {};
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                        __builtin_bit_cast(bf16x8, a[i]), __builtin_bit_cast(bf16x8, b[j]), acc[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) s += acc[i][j][0] + acc[i][j][15];
    } else if constexpr (FORM == 3) {
        float a[8], b[4];
#pragma unroll
        for (int i = 0; i < 8; ++i) a[i] = __uint_as_float(word(i)[0]);
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = __uint_as_float(word(8 + j)[1]);
        f4 acc[8][4] = {};
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) s += acc[i][j][0] + acc[i][j][3];
    } else if constexpr (FORM == 4) {
        float a[4], b[2];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = __uint_as_float(word(i)[0]);
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = __uint_as_float(word(8 + j)[1]);
        f16v acc[4][2] = {};
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], b[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) s += acc[i][j][0] + acc[i][j][15];
    } else if constexpr (FORM == 5) {
        double a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = (double)__uint_as_float(word(i)[0]);
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = (double)__uint_as_float(word(8 + j)[1]);
        d4 acc[4][4] = {};
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) s += (float)(acc[i][j][0] + acc[i][j][3]);
    } else if constexpr (FORM == 7) {
        i32x8 a[2], b[4];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const u32x4 x = word(i), y = word(i + 4);
            a[i] = i32x8{(int)x[0], (int)x[1], (int)x[2], (int)x[3], (int)y[0], (int)y[1], (int)y[2], (int)y[3]};
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const u32x4 x = word(8 + j), y = word(2 + j);
            b[j] = i32x8{(int)x[0], (int)x[1], (int)x[2], (int)x[3], (int)y[0], (int)y[1], (int)y[2], (int)y[3]};
        }
        f16v acc[2][4] = {};
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a[i], b[j], acc[i][j], 0, 0, 0, 127, 0,
                                                                                127);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) s += acc[i][j][0] + acc[i][j][15];
    } else {
        i32x8 a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const u32x4 x = word(i), y = word(i + 4);
            a[i] = i32x8{(int)x[0], (int)x[1], (int)x[2], (int)x[3], (int)y[0], (int)y[1], (int)y[2], (int)y[3]};
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const u32x4 x = word(8 + j), y = word(2 + j);
            b[j] = i32x8{(int)x[0], (int)x[1], (int)x[2], (int)x[3], (int)y[0], (int)y[1], (int)y[2], (int)y[3]};
        }
        f4 acc[4][4] = {};
        for (int it = 0; it < iters; ++it)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a[i], b[j], acc[i][j], 0, 0, 0, 127, 0,
                                                                                 127);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) s += acc[i][j][0] + acc[i][j][3];
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

constexpr int kLdsRegion = 2 << 20;   // bytes read by the load probe, wrapping
constexpr int kLdsStage = 32 << 10;   // bytes per workgroup per iteration
constexpr int kLdsPerWave = kLdsStage / 4 / 1024;  // 1 KiB loads per wave per iteration

// Each 4-wave workgroup loads kLdsStage bytes per iteration from a 2 MiB
// region into one of two LDS buffers, the previous iteration's loads still
// in flight.  MODE 0: direct-to-LDS buffer loads; 1: buffer loads into
// registers, then ds_write; 2: buffer loads into registers only.
template <int MODE>
__global__ void __launch_bounds__(256) lds_probe_kernel(const uint8_t *__restrict__ src, uint32_t *__restrict__ out,
                                                        int iters)
{
    __shared__ __attribute__((aligned(1024))) uint8_t lds[2][kLdsStage];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t acc = 0;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t *>(src), (short)0, kLdsRegion, 0x00020000);
    const uint32_t wg_off = (uint32_t)((blockIdx.x * 7919u) % (kLdsRegion / kLdsStage)) * kLdsStage;
    for (int it = 0; it < iters; ++it) {
        const int st = it & 1;
        const uint32_t base = (wg_off + (uint32_t)it * kLdsStage) % kLdsRegion;
        if constexpr (MODE == 0) {
#pragma unroll
            for (int q = 0; q < kLdsPerWave; ++q) {
                const int b = wave * kLdsPerWave + q;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rs, (__attribute__((address_space(3))) void *)(&lds[st][b * 1024]), 16, lane * 16, base + b * 1024, 0, 0);
            }
            // the previous iteration's loads have landed (vmcnt(kLdsPerWave))
            __builtin_amdgcn_s_waitcnt((kLdsPerWave & 15) | ((kLdsPerWave >> 4) << 14) | (7 << 4) | (15 << 8));
            __builtin_amdgcn_s_barrier();
        } else {
            u32x4 r[kLdsPerWave];
#pragma unroll
            for (int q = 0; q < kLdsPerWave; ++q)
                r[q] = __builtin_bit_cast(
                    u32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, base + (wave * kLdsPerWave + q) * 1024 + lane * 16, 0, 0));
            if constexpr (MODE == 1) {
#pragma unroll
                for (int q = 0; q < kLdsPerWave; ++q)
                    *reinterpret_cast<u32x4 *>(&lds[st][(wave * kLdsPerWave + q) * 1024 + lane * 16]) = r[q];
                __syncthreads();
            } else {
#pragma unroll
                for (int q = 0; q < kLdsPerWave; ++q) acc += r[q][0] ^ r[q][3];
            }
        }
    }
    __builtin_amdgcn_s_waitcnt(0);  // no load into LDS in flight at exit
    __syncthreads();
    if (MODE != 2) acc += lds[0][tid * 4] + lds[1][tid * 4];
    out[blockIdx.x * 256 + tid] = acc;
}

}  // namespace
}  // namespace blasdsp

extern "C" int blasdsp_probe_mfma(int form, const void *src, void *out, int blocks, int iters, void *stream)
{
    using namespace blasdsp;
    if (!src || !out || blocks <= 0 || iters <= 0 || ((uintptr_t)src % 16) != 0) return BLASDSP_ERR_INVALID;
    const hipStream_t s = (hipStream_t)stream;
    const u32x4 *in = (const u32x4 *)src;
    float *o = (float *)out;
    switch (form) {
    case 0: hipLaunchKernelGGL(mfma_probe_kernel<0>, dim3(blocks), dim3(256), 0, s, in, o, iters); break;
    case 1: hipLaunchKernelGGL(mfma_probe_kernel<1>, dim3(blocks), dim3(256), 0, s, in, o, iters); break;
    case 2: hipLaunchKernelGGL(mfma_probe_kernel<2>, dim3(blocks), dim3(256), 0, s, in, o, iters); break;
    case 3: hipLaunchKernelGGL(mfma_probe_kernel<3>, dim3(blocks), dim3(256), 0, s, in, o, iters); break;
    case 4: hipLaunchKernelGGL(mfma_probe_kernel<4>, dim3(blocks), dim3(256), 0, s, in, o, iters); break;
    case 5: hipLaunchKernelGGL(mfma_probe_kernel<5>, dim3(blocks), dim3(256), 0, s, in, o, iters); break;
    case 6: hipLaunchKernelGGL(mfma_probe_kernel<6>, dim3(blocks), dim3(256), 0, s, in, o, iters); break;
    case 7: hipLaunchKernelGGL(mfma_probe_kernel<7>, dim3(blocks), dim3(256), 0, s, in, o, iters); break;
    default: return BLASDSP_ERR_INVALID;
    }
    return hip_status(hipGetLastError());
}

extern "C" int blasdsp_probe_lds(int mode, const void *src, void *out, int blocks, int iters, void *stream)
{
    using namespace blasdsp;
    if (!src || !out || blocks <= 0 || iters <= 0) return BLASDSP_ERR_INVALID;
    const hipStream_t s = (hipStream_t)stream;
    const uint8_t *in = (const uint8_t *)src;
    uint32_t *o = (uint32_t *)out;
    switch (mode) {
    case 0: hipLaunchKernelGGL(lds_probe_kernel<0>, dim3(blocks), dim3(256), 0, s, in, o, iters); break;
    case 1: hipLaunchKernelGGL(lds_probe_kernel<1>, dim3(blocks), dim3(256), 0, s, in, o, iters); break;
    case 2: hipLaunchKernelGGL(lds_probe_kernel<2>, dim3(blocks), dim3(256), 0, s, in, o, iters); break;
    default: return BLASDSP_ERR_INVALID;
    }
    return hip_status(hipGetLastError());
}

extern "C" int blasdsp_probe_read(const void *p, int64_t bytes, void *out, int blocks, void *stream)
{
    if (!p || !out || bytes < 16 || bytes % 16 != 0 || blocks <= 0 || ((uintptr_t)p % 16) != 0)
        return BLASDSP_ERR_INVALID;
    hipLaunchKernelGGL(blasdsp::read_probe_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       (const blasdsp::u32x4 *)p, bytes / 16, (uint32_t *)out);
    return blasdsp::hip_status(hipGetLastError());
}
