// gemm_tile.hip — MFMA tile kernel for dgemm (v_mfma_f64_16x16x4_f64) and
// sgemm (v_mfma_f32_16x16x4_f32), all four transpose modes, any M/N/K >= 1 and
// any leading dimension.
//
// What it replaces.  The reference's dgemm_NN/NT/TN/TT and sgemm_*_qxx drive a
// Goto-style blocking with DMA double/triple buffering into the DSP's on-chip
// SM/AM/GSM and an outer-product FMAC micro-kernel (r6c64 / r6c128), plus
// on-chip 16x16 / 32x32 transposes or dedicated "transposer cores" for the T
// modes (dgemm_release_V1.1/device_code/dgemm_kernel.c:1994-3168,
// sgemm_release_V1.3.2/device_code/sgemm_kernel.c:2476-4500).  The per-tile
// work is tile::tile_body (tile_body.h); this file launches one workgroup per
// C tile.  In summary:
//   * one workgroup owns a 128 x 128 tile of C (64 x 64 when the 128 tiles
//     would not cover the CUs): 8 waves of 32 x 64 for the FP64 128 tile,
//     4 waves of quarter tiles otherwise;
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
#include <algorithm>
#include <cmath>

#include "tile_body.h"

namespace blasdsp {
namespace {

using tile::Cfg;

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

// Waves per workgroup of the per-call kernel: 8 for the FP64 128 tile (each
// wave 32 x 64, four waves per SIMD), 4 elsewhere.
template <typename T, int BM>
struct PerCall {
    static constexpr int NW = BM == 128 ? 8 : 4;
    using Cf = Cfg<T, BM, NW>;
};

// BODY: kMixed, the unchecked body for interior tiles whose K range is a
// whole number of K tiles and the checked body elsewhere; kFull, every tile
// takes the unchecked body (the host has checked that each does); kPtr,
// 64-bit operand addressing, for operands the buffer loads do not take
// (tile::buffer_ok).
enum { kMixed = 0, kFull = 1, kPtr = 2 };
template <typename T, int BM, bool TA, bool TB, int BODY>
__global__ void __launch_bounds__((PerCall<T, BM>::Cf::THREADS))
__attribute__((amdgpu_waves_per_eu(PerCall<T, BM>::Cf::WPE, PerCall<T, BM>::Cf::WPE)))
tile_gemm_kernel(Gemm<T> g, int tiles_m, int tiles_n, int64_t kslice, T *W)
{
    using Cf = typename PerCall<T, BM>::Cf;
    __shared__ __attribute__((aligned(16))) T lds[Cf::LDS_ELEMS];
    int tm, tn;
    tile_of_block(tiles_m, tiles_n, tm, tn);
    const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BM;
    // split-K: slice blockIdx.y of K writes its partial to W (one M x N per slice)
    const int64_t kb = (int64_t)blockIdx.y * kslice, ke = min(g.K, kb + kslice);
    T *w = W ? W + (int64_t)blockIdx.y * g.M * g.N : nullptr;
    if (BODY == kPtr)
        tile::tile_body<T, BM, TA, TB, Cf::NW, false, true>(g, m0, n0, lds, kb, ke, w);
    else if (BODY == kFull || tile::full_tile<Cf>(g, m0, n0, kb, ke))
        tile::tile_body<T, BM, TA, TB, Cf::NW, true>(g, m0, n0, lds, kb, ke, w);
    else
        tile::tile_body<T, BM, TA, TB, Cf::NW, false>(g, m0, n0, lds, kb, ke, w);
}

// C = alpha * sum_s W[s] + beta * C, slices added in order s = 0, 1, ...
// (deterministic); C is not read when beta == 0.
template <typename T>
__global__ void __launch_bounds__(256) tile_reduce_kernel(const T *__restrict__ W, int splitk, T *__restrict__ C,
                                                          int64_t M, int64_t N, int64_t ldc, T alpha, T beta)
{
    const int64_t total = M * N;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        T sum = W[e];
        for (int q = 1; q < splitk; ++q) sum += W[(int64_t)q * total + e];
        T *p = C + (e / N) * ldc + e % N;
        T v = alpha * sum;
        if (beta != T(0)) v += beta * *p;
        *p = v;
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

// K per split-K slice: a whole number of K tiles.
template <typename T>
int64_t tile_kslice(int64_t K, int splitk)
{
    constexpr int64_t BK = tile::Cfg<T, 128>::BK;
    const int64_t per = (K + splitk - 1) / splitk;
    return (per + BK - 1) / BK * BK;
}

template <typename T, int BM>
hipError_t launch_tile_bm(const Gemm<T> &g, int splitk, T *work, hipStream_t s)
{
    const int64_t tiles_m = (g.M + BM - 1) / BM;
    const int64_t tiles_n = (g.N + BM - 1) / BM;
    const int64_t nwg = tiles_m * tiles_n;
    if (nwg <= 0) return hipSuccess;
    if (nwg > 0x7fffffffLL) return hipErrorInvalidValue;
    const int64_t kslice = splitk > 1 ? tile_kslice<T>(g.K, splitk) : g.K;
    const int64_t slices = splitk > 1 ? (g.K + kslice - 1) / kslice : 1;
    if (slices > 1 && !work) return hipErrorInvalidValue;
    T *W = slices > 1 ? work : nullptr;
    const bool ptr = !tile::buffer_ok<T, BM>(g);
    const dim3 grid((unsigned)nwg, (unsigned)slices), block(PerCall<T, BM>::Cf::THREADS);
    const int tmi = (int)tiles_m, tni = (int)tiles_n;
    // every tile inside C and every K slice a whole number of K tiles
    constexpr int64_t BK = PerCall<T, BM>::Cf::BK;
    const bool all_full = g.M % BM == 0 && g.N % BM == 0 && kslice % BK == 0 && g.K % BK == 0;
    const int body = ptr ? kPtr : all_full ? kFull : kMixed;
    const int mode = (g.ta ? 2 : 0) | (g.tb ? 1 : 0);
    using K = void (*)(Gemm<T>, int, int, int64_t, T *);
    static constexpr K kernels[3][4] = {
        {tile_gemm_kernel<T, BM, false, false, kMixed>, tile_gemm_kernel<T, BM, false, true, kMixed>,
         tile_gemm_kernel<T, BM, true, false, kMixed>, tile_gemm_kernel<T, BM, true, true, kMixed>},
        {tile_gemm_kernel<T, BM, false, false, kFull>, tile_gemm_kernel<T, BM, false, true, kFull>,
         tile_gemm_kernel<T, BM, true, false, kFull>, tile_gemm_kernel<T, BM, true, true, kFull>},
        {tile_gemm_kernel<T, BM, false, false, kPtr>, tile_gemm_kernel<T, BM, false, true, kPtr>,
         tile_gemm_kernel<T, BM, true, false, kPtr>, tile_gemm_kernel<T, BM, true, true, kPtr>}};
    hipLaunchKernelGGL(kernels[body][mode], grid, block, 0, s, g, tmi, tni, kslice, W);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || slices == 1) return e;
    const int64_t total = g.M * g.N;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL((tile_reduce_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, s, W, (int)slices, g.C, g.M,
                       g.N, g.ldc, g.alpha, g.beta);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_tile(const Gemm<T> &g, int bm, int splitk, T *work, hipStream_t s)
{
    if (bm == 64) return launch_tile_bm<T, 64>(g, splitk, work, s);
    if (bm == 128) return launch_tile_bm<T, 128>(g, splitk, work, s);
    return hipErrorInvalidValue;
}

// Split-K for the tile kernel: the slice count 1 .. 16 (each slice at least
// 16 K tiles deep) that minimises the estimated time
//   ceil(units / slots) rounds x (one unit's time at full occupancy)
//   + the reduction's HBM traffic ((2 S + 2) M N elements at 5 TB/s),
// where units = C tiles x slices and slots = resident workgroups per CU x CUs.
// A round is charged in full even when partly filled, so a GEMM whose tiles
// already fill the slots keeps one slice.  Full-occupancy rates: FP64
// 65 TFLOP/s, FP32 130 TFLOP/s (the tile kernel at 4096^3 / 8192^3).
template <typename T>
int tile_pick_splitk(const Gemm<T> &g, int bm, int num_cus)
{
    const int64_t tiles = ((g.M + bm - 1) / bm) * ((g.N + bm - 1) / bm);
    const int64_t slots = (int64_t)(bm == 128 ? 2 : 4) * num_cus;
    constexpr int64_t BK = tile::Cfg<T, 128>::BK;
    const int64_t max_s = std::min<int64_t>(16, g.K / (16 * BK));
    if (tiles <= 0 || max_s < 2) return 1;
    const double rate = sizeof(T) == 8 ? 65e12 : 130e12, bw = 5e12;
    const double unit_full = 2.0 * bm * bm * (double)g.K * (double)slots / rate;  // one slice = K
    const double mn_bytes = (double)g.M * (double)g.N * sizeof(T);
    int best = 1;
    double best_t = std::ceil((double)tiles / (double)slots) * unit_full;
    for (int64_t sk = 2; sk <= max_s; ++sk) {
        const int64_t ks = tile_kslice<T>(g.K, (int)sk);
        const int64_t slices = (g.K + ks - 1) / ks;
        const double rounds = std::ceil((double)(tiles * slices) / (double)slots);
        const double t = rounds * unit_full * (double)ks / (double)g.K + (2.0 * slices + 2.0) * mn_bytes / bw;
        if (t < best_t * 0.97) {
            best_t = t;
            best = (int)slices;
        }
    }
    return best;
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

template hipError_t launch_tile<double>(const Gemm<double> &, int, int, double *, hipStream_t);
template hipError_t launch_tile<float>(const Gemm<float> &, int, int, float *, hipStream_t);
template int tile_pick_splitk<double>(const Gemm<double> &, int, int);
template int tile_pick_splitk<float>(const Gemm<float> &, int, int);
template hipError_t launch_scale<double>(const Gemm<double> &, hipStream_t);
template hipError_t launch_scale<float>(const Gemm<float> &, hipStream_t);

}  // namespace blasdsp
