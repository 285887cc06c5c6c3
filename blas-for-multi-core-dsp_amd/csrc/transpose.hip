// transpose.hip — out-of-place matrix transpose B = A^T (FP64 and FP32).
//
// What it replaces.  The reference's stand-alone transpose operator
// (transpose/device_code/kernel.c:1167 transpose_16x16_dbuf for FP64, :2002
// transpose_float_32x32_dbuf for FP32, launched by transpose/host_code/
// transpose.c:25-166) and the exported transpose_A2B / transpose_A2B_float32
// (dgemm_release_V1.1/include/dgemm_kernel.h:8, sgemm_release_V1.3.2/include/
// sgemm_kernel.h:8).  There each DSP core DMAs 128 x 128 blocks into its
// vector memory, transposes them with 16 x 16 (or 32 x 32) register
// micro-kernels and DMAs them back.
//
// Here one workgroup moves a 128 x 128 tile (64 x 64 for FP32 rows that are
// only 8-byte aligned and for unaligned operands): 16-byte loads along the
// rows of A (8-byte for those FP32 rows), scalar writes into a padded LDS
// tile at the transposed position, then stores of the same width along the
// rows of B.  Every element crosses HBM once each way; the kernel is
// HBM-bound (algorithmic bytes = 2 * M * N * sizeof(T)).  Edge tiles and
// unaligned operands take element-wise accesses.
#include "api_common.h"

namespace blasdsp {
namespace {

// VW elements per access in full tiles: 16 bytes when both operands allow it
// (rows 16-byte aligned), 8 bytes for FP32 with rows 8-byte aligned, else 1
// (element-wise, with the 16-byte lane mapping).
template <typename T, int VW>
struct TAcc {
    static constexpr int V = VW > 1 ? VW : 16 / (int)sizeof(T);  // elements per lane and row piece
    typedef T v __attribute__((ext_vector_type(V)));
};

// Tile shape and workgroup size per access width, measured on MI355X
// (interleaved A/B over 16384^2, the reference harness's 8210 x 8190 and
// 4096 x 8192; all exact):
//   * FP64, 16-byte rows: 128 x 128 on 1024 threads (one workgroup per CU by
//     LDS).  Against 64 x 64 on 256: 16384^2 5.21 -> 5.55-5.57 TB/s, 8210 x
//     8190 4.00 -> 4.75-4.77, 4096 x 8192 4.45 -> 5.21-5.27; on 512 threads
//     it ran 15 % slower, 64 x 128 on 512 gained about half as much;
//   * FP32, 16-byte rows: 128 x 128 on 512 threads (512-byte row pieces each
//     way): 16384^2 4.15 (64 x 64) -> 4.39 (on 256 threads) -> 5.15-5.21;
//     on 1024 threads 4.66;
//   * FP32, 8-byte rows: 64 x 64 on 512 threads: 8210 x 8190 3.03
//     (element-wise) -> 3.28 (on 256) -> 3.36-3.39; a 128 tile ran 10-20 %
//     slower;
//   * element-wise: 64 x 64 on 256 threads.
template <typename T, int VW>
constexpr bool big_tile()
{
    return (sizeof(T) == 8 && VW == 2) || (sizeof(T) == 4 && VW == 4);
}
template <typename T, int VW>
constexpr int tile_rows()
{
    return big_tile<T, VW>() ? 128 : 64;
}
template <typename T, int VW>
constexpr int tile_cols()
{
    return big_tile<T, VW>() ? 128 : 64;
}
template <typename T, int VW>
constexpr int tile_threads()
{
    return sizeof(T) == 8 && VW == 2 ? 1024 : (VW > 1 ? 512 : 256);
}

// A: M x N (lda), B: N x M (ldb).  Tile (by, bx) covers rows by*TR.. of A and
// columns bx*TC.. (TR = tile_rows, TC = tile_cols).
template <typename T, int VW>
__global__ void __launch_bounds__((tile_threads<T, VW>()))
transpose_kernel(const T *__restrict__ A, int64_t lda, T *__restrict__ B, int64_t ldb, int64_t M, int64_t N)
{
    using vec = typename TAcc<T, VW>::v;
    constexpr int V = TAcc<T, VW>::V, TR = tile_rows<T, VW>(), TC = tile_cols<T, VW>();
    constexpr int NT = tile_threads<T, VW>();
    constexpr int VPR = TC / V, VPC = TR / V;  // vectors per tile row (load), per tile column (store)
    static_assert(TR * VPR % NT == 0 && TC * VPC % NT == 0, "whole vectors per thread");
    __shared__ T tile[TC][TR + 1];
    const int tid = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.y * TR, c0 = (int64_t)blockIdx.x * TC;
    const bool full = VW > 1 && r0 + TR <= M && c0 + TC <= N;
    // load: thread (row, vector) = (e / VPR, e % VPR), e = tid + NT i
#pragma unroll
    for (int i = 0; i < TR * VPR / NT; ++i) {
        const int e = tid + NT * i, r = e / VPR, cv = (e % VPR) * V;
        vec x;
        if (full) {
            x = __builtin_nontemporal_load(reinterpret_cast<const vec *>(A + (r0 + r) * lda + c0 + cv));
        } else {
#pragma unroll
            for (int q = 0; q < V; ++q)
                x[q] = (r0 + r < M && c0 + cv + q < N) ? A[(r0 + r) * lda + c0 + cv + q] : T(0);
        }
#pragma unroll
        for (int q = 0; q < V; ++q) tile[cv + q][r] = x[q];
    }
    __syncthreads();
    // store: row c0 + rr of B holds column c0 + rr of A
#pragma unroll
    for (int i = 0; i < TC * VPC / NT; ++i) {
        const int e = tid + NT * i, rr = e / VPC, mv = (e % VPC) * V;
        vec x;
#pragma unroll
        for (int q = 0; q < V; ++q) x[q] = tile[rr][mv + q];
        if (full) {
            __builtin_nontemporal_store(x, reinterpret_cast<vec *>(B + (c0 + rr) * ldb + r0 + mv));
        } else {
#pragma unroll
            for (int q = 0; q < V; ++q)
                if (c0 + rr < N && r0 + mv + q < M) B[(c0 + rr) * ldb + r0 + mv + q] = x[q];
        }
    }
}

__global__ void transpose_null_kernel() {}

template <typename T>
int transpose(int64_t M, int64_t N, const T *A, int64_t lda, T *B, int64_t ldb, void *stream)
{
    if (M < 0 || N < 0) return BLASDSP_ERR_INVALID;
    if (M == 0 || N == 0) return BLASDSP_SUCCESS;
    if (!A || !B || lda < N || ldb < M) return BLASDSP_ERR_INVALID;
    // in place is not supported: the output must not overlap the input
    const uintptr_t a0 = (uintptr_t)A, a1 = (uintptr_t)(A + (M - 1) * lda + N);
    const uintptr_t b0 = (uintptr_t)B, b1 = (uintptr_t)(B + (N - 1) * ldb + M);
    if (a0 < b1 && b0 < a1) return BLASDSP_ERR_INVALID;
    constexpr int V16 = 16 / (int)sizeof(T);
    const bool vec16 = a0 % 16 == 0 && b0 % 16 == 0 && lda % V16 == 0 && ldb % V16 == 0;
    const bool vec8 = sizeof(T) == 4 && a0 % 8 == 0 && b0 % 8 == 0 && lda % 2 == 0 && ldb % 2 == 0;
    const int vw = vec16 ? V16 : (vec8 ? 2 : 1);
    const int64_t tr = vw == V16 ? tile_rows<T, V16>() : 64, tc = vw == V16 ? tile_cols<T, V16>() : 64;
    static_assert((V16 == 2 || tile_rows<T, 2>() == 64) && tile_rows<T, 1>() == 64, "8-byte and element-wise tiles are 64");
    const int64_t gx = (N + tc - 1) / tc, gy = (M + tr - 1) / tr;
    if (gx > 0x7fffffffLL || gy > 65535) return BLASDSP_ERR_INVALID;
    const dim3 grid((unsigned)gx, (unsigned)gy);
    const hipStream_t s = (hipStream_t)stream;
    if (vw == V16)
        hipLaunchKernelGGL((transpose_kernel<T, V16>), grid, dim3(tile_threads<T, V16>()), 0, s, A, lda, B, ldb, M, N);
    else if (vw == 2)
        hipLaunchKernelGGL((transpose_kernel<T, 2>), grid, dim3(tile_threads<T, 2>()), 0, s, A, lda, B, ldb, M, N);
    else
        hipLaunchKernelGGL((transpose_kernel<T, 1>), grid, dim3(tile_threads<T, 1>()), 0, s, A, lda, B, ldb, M, N);
    return hip_status(hipGetLastError());
}

int transpose_null(int64_t M, int64_t N, const void *A, const void *B, void *stream)
{
    if (M < 0 || N < 0 || (M > 0 && N > 0 && (!A || !B))) return BLASDSP_ERR_INVALID;
    hipLaunchKernelGGL(transpose_null_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream);
    return hip_status(hipGetLastError());
}

}  // namespace
}  // namespace blasdsp

extern "C" {

int blasdsp_transpose_d(int64_t M, int64_t N, const double *A, int64_t lda, double *B, int64_t ldb, void *stream)
{
    return blasdsp::transpose<double>(M, N, A, lda, B, ldb, stream);
}

int blasdsp_transpose_s(int64_t M, int64_t N, const float *A, int64_t lda, float *B, int64_t ldb, void *stream)
{
    return blasdsp::transpose<float>(M, N, A, lda, B, ldb, stream);
}

int transpose_A2B(int64_t lda, int64_t ldb, int64_t M, int64_t N, const double *A, double *B, void *stream)
{
    return blasdsp::transpose<double>(M, N, A, lda, B, ldb, stream);
}

int transpose_A2B_float32(int64_t M, int64_t N, const float *A, float *B, void *stream)
{
    return blasdsp::transpose<float>(M, N, A, N, B, M, stream);
}

int transpose_16x16_dbuf(int64_t M, int64_t N, const double *A, double *B, void *stream)
{
    return blasdsp::transpose<double>(M, N, A, N, B, M, stream);
}

int transpose_float_32x32_dbuf(int64_t M, int64_t N, const float *A, float *B, void *stream)
{
    return blasdsp::transpose<float>(M, N, A, N, B, M, stream);
}

int transpose_null(int64_t M, int64_t N, const double *A, double *B, void *stream)
{
    return blasdsp::transpose_null(M, N, A, B, stream);
}

int transpose_float_null(int64_t M, int64_t N, const float *A, float *B, void *stream)
{
    return blasdsp::transpose_null(M, N, A, B, stream);
}

}  // extern "C"
