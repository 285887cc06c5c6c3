// gemm_tile.hip — MFMA tile kernel for dgemm (v_mfma_f64_16x16x4_f64) and
// sgemm (v_mfma_f32_16x16x4_f32), all four transpose modes, any M/N/K >= 1 and
// any leading dimension.
//
// What it replaces.  The reference's dgemm_NN/NT/TN/TT and sgemm_*_qxx drive a
// Goto-style blocking with DMA double/triple buffering into the DSP's on-chip
// SM/AM/GSM and an outer-product FMAC micro-kernel (r6c64 / r6c128), plus
// on-chip 16x16 / 32x32 transposes or dedicated "transposer cores" for the T
// modes (dgemm_release_V1.1/device_code/dgemm_kernel.c:1994-3168,
// sgemm_release_V1.3.2/device_code/sgemm_kernel.c:2476-4500).  Here:
//   * one 256-thread workgroup owns a 128 x 128 tile of C; its four waves own
//     64 x 64 sub-tiles, i.e. 4 x 4 MFMA 16x16 accumulators each;
//   * op(A) / op(B) tiles of BK = 128 bytes of K are staged global -> LDS with
//     16-byte vector loads along whichever dimension is contiguous in HBM, and
//     the LDS copy keeps that same orientation.  The MFMA fragment reads adapt
//     to the orientation instead, so the transpose modes cost nothing: no
//     transpose pass and no transposing LDS writes;
//   * LDS row strides are padded (k-contiguous rows by 16 B, m/n-contiguous
//     rows by 16 elements) so that every fragment read is bank-conflict-free
//     for a half-wave;
//   * two LDS stages: the global loads for tile k+1 are in flight while the
//     MFMAs consume tile k, one barrier per K tile;
//   * the workgroup -> tile map is XCD-aware (the dispatcher places block b on
//     XCD b % 8) and grouped by 8 tile rows so that concurrently running tiles
//     share A and B panels in each XCD's L2;
//   * epilogue C = alpha * acc + beta * C fused into the store (C not read
//     when beta == 0).
#include "internal.h"

namespace blasdsp {
namespace {

template <typename T>
struct Mma;

template <>
struct Mma<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    typedef double vec_t __attribute__((ext_vector_type(2)));
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // Output layout of the f64 16x16 block: lane group g = lane / 16 and
    // register r hold row g + 4 r (column lane % 16).
    static __device__ __forceinline__ int out_row(int g, int r) { return g + 4 * r; }
};

template <>
struct Mma<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    typedef float vec_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // f32 16x16 block: lane group g, register r hold row 4 g + r.
    static __device__ __forceinline__ int out_row(int g, int r) { return 4 * g + r; }
};

template <typename T>
struct Cfg {
    static constexpr int BM = 128, BN = 128;
    static constexpr int BK = 128 / (int)sizeof(T);  // one 128-byte row of K
    static constexpr int V = 16 / (int)sizeof(T);    // elements per 16-byte vector
    static constexpr int SK = BK + V;                // k-contiguous LDS row stride
    static constexpr int SMN = BM + 16;              // m/n-contiguous LDS row stride
    static constexpr int BUF = BM * SK;              // elements per operand stage
    static constexpr int THREADS = 256;
    static constexpr int WM = 64, WN = 64;           // wave sub-tile
    static constexpr int MT = WM / 16, NT = WN / 16; // MFMA tiles per wave
    static_assert(BM == BN, "square workgroup tile");
    static_assert(BM * SK == BK * SMN, "both LDS orientations have the same footprint");
};

// Global -> register -> LDS staging of one operand tile.  KC: the stored tile is
// k-contiguous (BM rows of BK elements); otherwise m/n-contiguous (BK rows of
// BM elements).  Either way it is 1024 16-byte vectors, 4 per thread; a wave
// reads 8 full 128-byte rows (KC) or one full 1 KiB row (!KC) per instruction.
template <typename T, bool KC>
struct Stager {
    using Cf = Cfg<T>;
    using vec_t = typename Mma<T>::vec_t;
    static constexpr int TR = KC ? Cf::BM : Cf::BK;
    static constexpr int TC = KC ? Cf::BK : Cf::BM;
    static constexpr int CPR = TC / Cf::V;
    static constexpr int NCH = TR * CPR / Cf::THREADS;
    static constexpr int S = KC ? Cf::SK : Cf::SMN;
    static_assert(TR * CPR % Cf::THREADS == 0, "whole vectors per thread");

    vec_t r[NCH];

    // X: stored matrix with R rows, Cn columns, leading dimension ld; the tile
    // starts at stored (r0, c0).  Out-of-range elements read as zero.
    __device__ __forceinline__ void load(const T *__restrict__ X, int64_t ld, int64_t R, int64_t Cn,
                                         int64_t r0, int64_t c0, bool vec, int tid)
    {
        const bool full = vec && (r0 + TR <= R) && (c0 + TC <= Cn);
        if (full) {
#pragma unroll
            for (int i = 0; i < NCH; ++i) {
                const int c = tid + i * Cf::THREADS;
                const int row = c / CPR, col = (c % CPR) * Cf::V;
                r[i] = *reinterpret_cast<const vec_t *>(X + (r0 + row) * ld + c0 + col);
            }
        } else {
#pragma unroll
            for (int i = 0; i < NCH; ++i) {
                const int c = tid + i * Cf::THREADS;
                const int row = c / CPR, col = (c % CPR) * Cf::V;
                const int64_t gr = r0 + row, gc = c0 + col;
                if (vec && gr < R && gc + Cf::V <= Cn) {
                    r[i] = *reinterpret_cast<const vec_t *>(X + gr * ld + gc);
                } else {
#pragma unroll
                    for (int e = 0; e < Cf::V; ++e)
                        r[i][e] = (gr < R && gc + e < Cn) ? X[gr * ld + gc + e] : T(0);
                }
            }
        }
    }

    __device__ __forceinline__ void store(T *s, int tid) const
    {
#pragma unroll
        for (int i = 0; i < NCH; ++i) {
            const int c = tid + i * Cf::THREADS;
            const int row = c / CPR, col = (c % CPR) * Cf::V;
            *reinterpret_cast<vec_t *>(s + row * S + col) = r[i];
        }
    }
};

// Block -> C tile.  The hardware dispatcher places workgroup b on XCD b % 8;
// renumbering gives each XCD a contiguous run of logical tile ids, and the
// grouped order walks GROUP tile-rows at a time so the tiles resident on one
// XCD share A row-panels and B column-panels in that XCD's L2.
__device__ __forceinline__ void tile_of_block(int tiles_m, int tiles_n, int &tm, int &tn)
{
    constexpr int GROUP = 8;
    int pid = blockIdx.x;
    const int nwg = gridDim.x;
    if ((nwg & 7) == 0) pid = (pid & 7) * (nwg >> 3) + (pid >> 3);
    const int width = GROUP * tiles_n;
    const int first_m = (pid / width) * GROUP;
    const int gsz = min(tiles_m - first_m, GROUP);
    tm = first_m + (pid % width) % gsz;
    tn = (pid % width) / gsz;
}

template <typename T, bool TA, bool TB>
__global__ void __launch_bounds__(256) tile_gemm_kernel(Gemm<T> g, int tiles_m, int tiles_n,
                                                        bool a_vec, bool b_vec)
{
    using Cf = Cfg<T>;
    using acc_t = typename Mma<T>::acc_t;
    constexpr int BK = Cf::BK, SK = Cf::SK, SMN = Cf::SMN;

    __shared__ __attribute__((aligned(16))) T lds[2][2][Cf::BUF];  // [stage][A, B]

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int wm = (wave >> 1) * Cf::WM, wn = (wave & 1) * Cf::WN;

    int tm, tn;
    tile_of_block(tiles_m, tiles_n, tm, tn);
    const int64_t m0 = (int64_t)tm * Cf::BM, n0 = (int64_t)tn * Cf::BN;

    // Stored geometry: A is K x M when transposed, else M x K; B is N x K when
    // transposed, else K x N.
    const int64_t AR = TA ? g.K : g.M, AC = TA ? g.M : g.K;
    const int64_t BR = TB ? g.N : g.K, BC = TB ? g.K : g.N;

    Stager<T, !TA> sa;
    Stager<T, TB> sb;
    auto load_tiles = [&](int64_t k0) {
        if (TA) sa.load(g.A, g.lda, AR, AC, k0, m0, a_vec, tid);
        else    sa.load(g.A, g.lda, AR, AC, m0, k0, a_vec, tid);
        if (TB) sb.load(g.B, g.ldb, BR, BC, n0, k0, b_vec, tid);
        else    sb.load(g.B, g.ldb, BR, BC, k0, n0, b_vec, tid);
    };

    acc_t acc[Cf::MT][Cf::NT];
#pragma unroll
    for (int i = 0; i < Cf::MT; ++i)
#pragma unroll
        for (int j = 0; j < Cf::NT; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    const int nk = (int)((g.K + BK - 1) / BK);
    load_tiles(0);
    sa.store(lds[0][0], tid);
    sb.store(lds[0][1], tid);
    __syncthreads();

    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) load_tiles((int64_t)(kt + 1) * BK);

        const T *As = lds[cur][0];
        const T *Bs = lds[cur][1];
#pragma unroll
        for (int ks = 0; ks < BK / 4; ++ks) {
            const int k = ks * 4 + lk;
            T a[Cf::MT], b[Cf::NT];
#pragma unroll
            for (int i = 0; i < Cf::MT; ++i) {
                const int m = wm + i * 16 + lr;
                a[i] = TA ? As[k * SMN + m] : As[m * SK + k];
            }
#pragma unroll
            for (int j = 0; j < Cf::NT; ++j) {
                const int n = wn + j * 16 + lr;
                b[j] = TB ? Bs[n * SK + k] : Bs[k * SMN + n];
            }
#pragma unroll
            for (int i = 0; i < Cf::MT; ++i)
#pragma unroll
                for (int j = 0; j < Cf::NT; ++j) acc[i][j] = Mma<T>::mma(a[i], b[j], acc[i][j]);
        }

        if (more) {
            sa.store(lds[cur ^ 1][0], tid);
            sb.store(lds[cur ^ 1][1], tid);
        }
        __syncthreads();
    }

    // MFMA 16x16 output: lane l, register r holds column l % 16 and row
    // Mma<T>::out_row(l / 16, r) of the 16 x 16 block.
    const T alpha = g.alpha, beta = g.beta;
#pragma unroll
    for (int i = 0; i < Cf::MT; ++i) {
#pragma unroll
        for (int j = 0; j < Cf::NT; ++j) {
            const int64_t col = n0 + wn + j * 16 + lr;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = m0 + wm + i * 16 + Mma<T>::out_row(lk, r);
                if (row < g.M && col < g.N) {
                    T *p = g.C + row * g.ldc + col;
                    T v = alpha * acc[i][j][r];
                    if (beta != T(0)) v += beta * *p;
                    *p = v;
                }
            }
        }
    }
}

template <typename T>
__global__ void scale_kernel(T *C, int64_t M, int64_t N, int64_t ldc, T beta)
{
    const int64_t total = M * N;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / N, j = e % N;
        T *p = C + i * ldc + j;
        *p = (beta == T(0)) ? T(0) : beta * *p;
    }
}

__global__ void null_kernel() {}

}  // namespace

template <typename T>
hipError_t launch_tile(const Gemm<T> &g, hipStream_t s)
{
    using Cf = Cfg<T>;
    const int64_t tiles_m = (g.M + Cf::BM - 1) / Cf::BM;
    const int64_t tiles_n = (g.N + Cf::BN - 1) / Cf::BN;
    const int64_t nwg = tiles_m * tiles_n;
    if (nwg <= 0) return hipSuccess;
    if (nwg > 0x7fffffffLL || g.K > 0x7fffffffLL * (int64_t)Cf::BK) return hipErrorInvalidValue;
    const bool a_vec = ((uintptr_t)g.A % 16 == 0) && (g.lda % Cf::V == 0);
    const bool b_vec = ((uintptr_t)g.B % 16 == 0) && (g.ldb % Cf::V == 0);
    const dim3 grid((unsigned)nwg), block(Cf::THREADS);
    const int tmi = (int)tiles_m, tni = (int)tiles_n;
    if (!g.ta && !g.tb)
        hipLaunchKernelGGL((tile_gemm_kernel<T, false, false>), grid, block, 0, s, g, tmi, tni, a_vec, b_vec);
    else if (!g.ta && g.tb)
        hipLaunchKernelGGL((tile_gemm_kernel<T, false, true>), grid, block, 0, s, g, tmi, tni, a_vec, b_vec);
    else if (g.ta && !g.tb)
        hipLaunchKernelGGL((tile_gemm_kernel<T, true, false>), grid, block, 0, s, g, tmi, tni, a_vec, b_vec);
    else
        hipLaunchKernelGGL((tile_gemm_kernel<T, true, true>), grid, block, 0, s, g, tmi, tni, a_vec, b_vec);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_scale(const Gemm<T> &g, hipStream_t s)
{
    const int64_t total = g.M * g.N;
    if (total <= 0) return hipSuccess;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL((scale_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, s, g.C, g.M, g.N,
                       g.ldc, g.beta);
    return hipGetLastError();
}

hipError_t launch_null(hipStream_t s)
{
    hipLaunchKernelGGL(null_kernel, dim3(1), dim3(64), 0, s);
    return hipGetLastError();
}

template hipError_t launch_tile<double>(const Gemm<double> &, hipStream_t);
template hipError_t launch_tile<float>(const Gemm<float> &, hipStream_t);
template hipError_t launch_scale<double>(const Gemm<double> &, hipStream_t);
template hipError_t launch_scale<float>(const Gemm<float> &, hipStream_t);

}  // namespace blasdsp
