// gemm_split.hip — FP32 GEMM on the 16-bit matrix cores ("split" forms).
//
// No counterpart in the reference (an added precision form).  Every FP32
// operand x is split into 16-bit pieces and the product is assembled from the
// piece products that matter:
//   BF16, 6 terms:  x = hi + mid + lo (each the round-to-nearest-even BF16 of
//                   what is left, residual ~2^-24 |x|);
//                   hh + hm + mh + hl + lh + mm (dropped terms ~2^-24)
//   BF16, 3 terms:  x = hi + lo (residual ~2^-17 |x|); hh + hl + lh
//   FP16, 3 terms:  rows of op(A) and columns of op(B) scaled by a power of two
//                   (exact) so the largest magnitude is in [2^14, 2^15);
//                   x' = hi + 2^-11 lo with 11-bit hi and lo (residual ~2^-22);
//                   hh + 2^-11 (hl + lh), the cross terms in their own
//                   accumulator, scales undone in the epilogue.
// Piece products are exact in FP32 (8- or 11-bit significands), so the error
// comes from the dropped terms, the split residual and FP32 accumulation.  The
// BF16 forms accumulate in two levels: each 256-deep K chunk from zero in its
// own MFMA accumulators, then added into the running sum.
//
// The six-product BF16 form runs on fragment-tiled pieces in
// gemm_split_frag.hip; the three-product forms run here.
//
// Kernels:
//   rowscale_kernel  (FP16 form) per-row power-of-two scales of op(A), op(B)^T;
//   split_kernel     reads op(A) (or op(B)^T) in any storage order and writes
//                    the pieces as planes [rows][Kp], K contiguous, zero padded
//                    to whole 128-row tiles and 64-deep K blocks — the
//                    transpose modes are absorbed here, through an LDS tile;
//   planes_kernel    a 128 x 128-tile MFMA GEMM (v_mfma_f32_16x16x32_bf16 /
//                    _f16) over the planes: per 32-deep K step each wave reads
//                    every piece fragment once from LDS and issues
//                    terms x 16 MFMAs.  The two-piece forms fit two
//                    workgroups per CU (2 x 80 KiB of LDS).
#include "internal.h"

namespace blasdsp {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kBM = 128, kBN = 128, kBK = 32;  // GEMM tile
constexpr int kKpad = 64;                      // plane K padding (split tile)
constexpr int kRowPad = 128;                   // plane row padding
constexpr int kLdsStride = kBK + 8;            // 16-bit elements per LDS row (80 bytes)
constexpr int kChunkTiles = 8;                 // K tiles per accumulation chunk (256)

// Round-to-nearest-even FP32 -> BF16 bits (finite inputs).
__device__ __forceinline__ uint32_t bf16_bits(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ float bf16_val(uint32_t h) { return __uint_as_float(h << 16); }

// Pieces of one FP32 value.  BF16: x = hi + mid + lo + r, each piece the
// round-to-nearest BF16 of what is left.  FP16 (scaled): with x' = x * 2^e
// (exact), hi = fp16(x'), lo = fp16((x' - hi) * 2^11); x' - hi is exact in
// FP32 and the pair carries 22 significant bits.
template <int NP, bool FP16>
__device__ __forceinline__ void split(float x, float scale, uint32_t (&p)[NP])
{
    if (FP16) {
        const float xs = x * scale;
        const _Float16 hi = (_Float16)xs;
        const float r = xs - (float)hi;
        const _Float16 lo = (_Float16)(r * 2048.0f);
        p[0] = __builtin_bit_cast(uint16_t, hi);
        if (NP > 1) p[NP > 1 ? 1 : 0] = __builtin_bit_cast(uint16_t, lo);
    } else {
        float r = x;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            p[i] = bf16_bits(r);
            r = r - bf16_val(p[i]);
        }
    }
}

// Power-of-two row scales for the FP16 split: 2^e with e chosen so that the
// row's largest magnitude lands in [2^14, 2^15) (below the FP16 maximum
// 65504); inv = 2^-e undoes it in the epilogue.
__device__ __forceinline__ void store_scale(float amax, int64_t r, float *scale, float *inv)
{
    int e = 0;
    if (amax > 0.0f && amax < INFINITY) {
        int p;
        frexpf(amax, &p);  // amax = m * 2^p, m in [0.5, 1)
        e = 15 - p;
        e = e < -120 ? -120 : (e > 120 ? 120 : e);
    }
    scale[r] = ldexpf(1.0f, e);
    inv[r] = ldexpf(1.0f, -e);
}

typedef float f4v __attribute__((ext_vector_type(4)));

// X stored [R][K]: one workgroup per row (rows R .. Rp-1 get scale 1).  VEC:
// 16-byte loads (X 16-byte aligned, ld a multiple of 4).
template <bool VEC>
__global__ void __launch_bounds__(256) rowscale_rows_kernel(const float *__restrict__ X, int64_t ld, int64_t R,
                                                            int64_t K, float *__restrict__ scale,
                                                            float *__restrict__ inv)
{
    __shared__ float part[256];
    const int64_t r = blockIdx.x;
    float amax = 0.0f;
    if (r < R) {
        const float *row = X + r * ld;
        int64_t k = 0;
        if (VEC) {
            for (k = 4 * (int64_t)threadIdx.x; k + 3 < K; k += 1024) {
                const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(row + k));
                amax = fmaxf(amax, fmaxf(fmaxf(fabsf(v[0]), fabsf(v[1])), fmaxf(fabsf(v[2]), fabsf(v[3]))));
            }
            k = K / 4 * 4 + threadIdx.x;  // the ragged end, one element per thread
            if (k < K) amax = fmaxf(amax, fabsf(row[k]));
        } else {
            for (k = threadIdx.x; k < K; k += 256) amax = fmaxf(amax, fabsf(row[k]));
        }
    }
    part[threadIdx.x] = amax;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] = fmaxf(part[threadIdx.x], part[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) store_scale(part[0], r, scale, inv);
}

// X stored [K][R] (row r of op(X) is column r): partial maxima of each column
// over K slices (blockIdx.y), coalesced across the columns of a workgroup.
__global__ void __launch_bounds__(256) colmax_kernel(const float *__restrict__ X, int64_t ld, int64_t R,
                                                     int64_t K, int64_t kslice, float *__restrict__ part,
                                                     int64_t Rp)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= Rp) return;
    const int64_t k0 = (int64_t)blockIdx.y * kslice, k1 = min(K, k0 + kslice);
    float amax = 0.0f;
    if (r < R)
        for (int64_t k = k0; k < k1; ++k) amax = fmaxf(amax, fabsf(X[k * ld + r]));
    part[(int64_t)blockIdx.y * Rp + r] = amax;
}

__global__ void __launch_bounds__(256) colscale_kernel(const float *__restrict__ part, int nslices, int64_t Rp,
                                                       float *__restrict__ scale, float *__restrict__ inv)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= Rp) return;
    float amax = 0.0f;
    for (int q = 0; q < nslices; ++q) amax = fmaxf(amax, part[(int64_t)q * Rp + r]);
    store_scale(amax, r, scale, inv);
}

// Pieces of four consecutive k of one row, written as one 8-byte store per
// plane (planes are [Rp][Kp] with Kp a multiple of 64: 8-byte aligned).
template <int NP, bool FP16>
__device__ __forceinline__ void split_store4(f4v x, float sc, uint16_t *__restrict__ planes, int64_t plane_elems,
                                             int64_t off)
{
    uint32_t p[4][NP];
#pragma unroll
    for (int j = 0; j < 4; ++j) split<NP, FP16>(x[j], sc, p[j]);
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        const uint2 w = make_uint2(p[0][q] | (p[1][q] << 16), p[2][q] | (p[3][q] << 16));
        *reinterpret_cast<uint2 *>(planes + q * plane_elems + off) = w;
    }
}

// One 64 x 64 tile of op(X) (R logical rows x K) into NP planes [Rp][Kp]:
// thread (row, k quad) = (e / 16, e % 16) for e = tid + 256 i.  KC: X is
// stored [R][K] (k contiguous), read as 16-byte vectors along k; else [K][R],
// read as 16-byte vectors along r into an LDS tile and taken back along k.
// VEC: 16-byte loads allowed (X 16-byte aligned, ld a multiple of 4).
template <int NP, bool FP16, bool KC, bool VEC>
__global__ void __launch_bounds__(256) split_kernel(const float *__restrict__ X, int64_t ld, int64_t R,
                                                   int64_t K, uint16_t *__restrict__ planes,
                                                   int64_t Kp, int64_t plane_elems,
                                                   const float *__restrict__ scale)
{
    __shared__ float tile[64][65];  // [r][k]
    const int tid = threadIdx.x;
    const int64_t k0 = (int64_t)blockIdx.x * 64, r0 = (int64_t)blockIdx.y * 64;
    if (!KC) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int e = tid + i * 256;
            const int kk = e / 16, rq = (e % 16) * 4;
            const int64_t gk = k0 + kk, gr = r0 + rq;
            f4v v = {0.0f, 0.0f, 0.0f, 0.0f};
            if (gk < K) {
                if (VEC && gr + 3 < R) {
                    v = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(X + gk * ld + gr));
                } else {
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        if (gr + c < R) v[c] = X[gk * ld + gr + c];
                }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) tile[rq + c][kk] = v[c];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int e = tid + i * 256;
        const int row = e / 16, kq = (e % 16) * 4;
        const int64_t gr = r0 + row, gk = k0 + kq;
        f4v x = {0.0f, 0.0f, 0.0f, 0.0f};
        if (KC) {
            if (gr < R) {
                if (VEC && gk + 3 < K) {
                    x = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(X + gr * ld + gk));
                } else {
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (gk + j < K) x[j] = X[gr * ld + gk + j];
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) x[j] = tile[row][kq + j];
        }
        split_store4<NP, FP16>(x, FP16 ? scale[gr] : 1.0f, planes, plane_elems, gr * Kp + gk);
    }
}

// Schemes: pieces per operand, products (piece of A, piece of B, accumulator),
// small terms first, whether accumulation is two-level, waves per SIMD.
enum { kBF16x6 = 0, kBF16x3 = 1, kFP16x3 = 2 };
template <int S>
struct Scheme;
template <>
struct Scheme<kBF16x3> {
    static constexpr int NP = 2, NT = 3, NACC = 1, WPE = 2;
    static constexpr bool FP16 = false, CHUNK = true;
    static constexpr int a[3] = {0, 1, 0};
    static constexpr int b[3] = {1, 0, 0};
    static constexpr int acc[3] = {0, 0, 0};
};
template <>
struct Scheme<kFP16x3> {
    // hi*hi into acc 0; the cross terms (carrying 2^11) into acc 1
    static constexpr int NP = 2, NT = 3, NACC = 2, WPE = 2;
    static constexpr bool FP16 = true, CHUNK = false;
    static constexpr int a[3] = {0, 1, 0};
    static constexpr int b[3] = {1, 0, 0};
    static constexpr int acc[3] = {1, 1, 0};
};

template <bool FP16>
__device__ __forceinline__ f4 mfma16(u32x4 a, u32x4 b, f4 c)
{
    if (FP16)
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b),
                                                      c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

// f32 16x16 output of the 16-bit MFMAs: lane group g, register r hold row 4 g + r.
__device__ __forceinline__ int out_row(int g, int r) { return 4 * g + r; }

template <int S>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Scheme<S>::WPE, Scheme<S>::WPE)))
planes_kernel(const uint16_t *__restrict__ Ap, const uint16_t *__restrict__ Bp, int64_t Kp, int64_t a_plane,
              int64_t b_plane, int tiles_m, int tiles_n, float *__restrict__ C, int64_t ldc, int64_t M,
              int64_t N, float alpha, float beta, const float *__restrict__ a_inv,
              const float *__restrict__ b_inv)
{
    using Sc = Scheme<S>;
    constexpr int NP = Sc::NP;
    constexpr int PLANE_LDS = kBM * kLdsStride;           // 16-bit elements per plane per stage
    __shared__ __attribute__((aligned(16))) uint16_t lds[2][2][NP][PLANE_LDS];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lg = lane >> 4;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;

    // XCD-aware grouped tile order (see gemm_tile.hip).
    int pid = blockIdx.x;
    const int nwg = gridDim.x;
    if ((nwg & 7) == 0) pid = (pid & 7) * (nwg >> 3) + (pid >> 3);
    const int width = 8 * tiles_n;
    const int first_m = (pid / width) * 8;
    const int gsz = min(tiles_m - first_m, 8);
    const int tm = first_m + (pid % width) % gsz, tn = (pid % width) / gsz;
    const int64_t m0 = (int64_t)tm * kBM, n0 = (int64_t)tn * kBN;

    // staging: per plane and operand 128 rows x 32 values = 512 16-byte vectors, 2 per thread
    u32x4 ra[NP][2], rb[NP][2];
    auto gload = [&](int64_t k0) {
#pragma unroll
        for (int p = 0; p < NP; ++p)
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int c = tid + i * 256, row = c >> 2, col = (c & 3) * 8;
                ra[p][i] = *reinterpret_cast<const u32x4 *>(Ap + p * a_plane + (m0 + row) * Kp + k0 + col);
                rb[p][i] = *reinterpret_cast<const u32x4 *>(Bp + p * b_plane + (n0 + row) * Kp + k0 + col);
            }
    };
    auto lstore = [&](int st) {
#pragma unroll
        for (int p = 0; p < NP; ++p)
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int c = tid + i * 256, row = c >> 2, col = (c & 3) * 8;
                *reinterpret_cast<u32x4 *>(&lds[st][0][p][row * kLdsStride + col]) = ra[p][i];
                *reinterpret_cast<u32x4 *>(&lds[st][1][p][row * kLdsStride + col]) = rb[p][i];
            }
    };

    f4 acc[Sc::NACC][4][4];
    f4 part[Sc::CHUNK ? 4 : 1][Sc::CHUNK ? 4 : 1];
#pragma unroll
    for (int q = 0; q < Sc::NACC; ++q)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[q][i][j] = f4{0, 0, 0, 0};
    if (Sc::CHUNK)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) part[Sc::CHUNK ? i : 0][Sc::CHUNK ? j : 0] = f4{0, 0, 0, 0};

    const int nk = (int)(Kp / kBK);
    gload(0);
    lstore(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) gload((int64_t)(kt + 1) * kBK);

        u32x4 fa[NP][4], fb[NP][4];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                fa[p][i] = *reinterpret_cast<const u32x4 *>(&lds[cur][0][p][(wm + i * 16 + lr) * kLdsStride + lg * 8]);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                fb[p][j] = *reinterpret_cast<const u32x4 *>(&lds[cur][1][p][(wn + j * 16 + lr) * kLdsStride + lg * 8]);
        }
        // Every accumulator sees its terms in the order t = 0, 1, ... in both
        // loop orders.  One accumulator per tile (two-level forms): terms
        // outermost, so consecutive MFMAs never share an accumulator.  Two
        // accumulators (FP16 form): tiles outermost, the three terms of a
        // tile back to back (measured 10 % faster at 8192^3).
        if (Sc::CHUNK) {
#pragma unroll
            for (int t = 0; t < Sc::NT; ++t)
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        part[Sc::CHUNK ? i : 0][Sc::CHUNK ? j : 0] = mfma16<Sc::FP16>(
                            fa[Sc::a[t]][i], fb[Sc::b[t]][j], part[Sc::CHUNK ? i : 0][Sc::CHUNK ? j : 0]);
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int t = 0; t < Sc::NT; ++t)
                        acc[Sc::acc[t]][i][j] = mfma16<Sc::FP16>(fa[Sc::a[t]][i], fb[Sc::b[t]][j], acc[Sc::acc[t]][i][j]);
        }
        if (Sc::CHUNK && ((kt + 1) % kChunkTiles == 0 || !more)) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    acc[0][i][j] += part[Sc::CHUNK ? i : 0][Sc::CHUNK ? j : 0];
                    part[Sc::CHUNK ? i : 0][Sc::CHUNK ? j : 0] = f4{0, 0, 0, 0};
                }
        }
        if (more) lstore(cur ^ 1);
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t col = n0 + wn + j * 16 + lr;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = m0 + wm + i * 16 + out_row(lg, r);
                if (row < M && col < N) {
                    float v = acc[0][i][j][r];
                    if (Sc::NACC > 1) v += 0x1.0p-11f * acc[Sc::NACC - 1][i][j][r];
                    if (Sc::FP16) v = v * a_inv[row] * b_inv[col];
                    float *p = C + row * ldc + col;
                    v *= alpha;
                    if (beta != 0.0f) v += beta * *p;
                    *p = v;
                }
            }
        }
}

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

template <int NP, bool FP16>
hipError_t launch_split(const float *X, int64_t ld, bool kc, int64_t R, int64_t K, uint16_t *planes,
                        int64_t Rp, int64_t Kp, const float *scale, hipStream_t s)
{
    if (Rp / 64 > 65535 || Kp / 64 > 0x7fffffffLL) return hipErrorInvalidValue;  // grid limits
    const dim3 grid((unsigned)(Kp / 64), (unsigned)(Rp / 64));
    const bool vec = ((uintptr_t)X % 16) == 0 && ld % 4 == 0;
    const int64_t pe = Rp * Kp;
    if (kc && vec)
        hipLaunchKernelGGL((split_kernel<NP, FP16, true, true>), grid, dim3(256), 0, s, X, ld, R, K, planes, Kp, pe, scale);
    else if (kc)
        hipLaunchKernelGGL((split_kernel<NP, FP16, true, false>), grid, dim3(256), 0, s, X, ld, R, K, planes, Kp, pe, scale);
    else if (vec)
        hipLaunchKernelGGL((split_kernel<NP, FP16, false, true>), grid, dim3(256), 0, s, X, ld, R, K, planes, Kp, pe, scale);
    else
        hipLaunchKernelGGL((split_kernel<NP, FP16, false, false>), grid, dim3(256), 0, s, X, ld, R, K, planes, Kp, pe, scale);
    return hipGetLastError();
}

constexpr int kScaleSlices = 64;  // K slices of the column-maximum pass

hipError_t launch_rowscale(const float *X, int64_t ld, bool kc, int64_t R, int64_t K, float *scale, float *inv,
                           int64_t Rp, float *part, hipStream_t s)
{
    if (kc) {
        if (((uintptr_t)X % 16) == 0 && ld % 4 == 0)
            hipLaunchKernelGGL(rowscale_rows_kernel<true>, dim3((unsigned)Rp), dim3(256), 0, s, X, ld, R, K, scale, inv);
        else
            hipLaunchKernelGGL(rowscale_rows_kernel<false>, dim3((unsigned)Rp), dim3(256), 0, s, X, ld, R, K, scale, inv);
        return hipGetLastError();
    }
    const int64_t kslice = (K + kScaleSlices - 1) / kScaleSlices;
    const int nslices = (int)((K + kslice - 1) / kslice);
    const unsigned bx = (unsigned)((Rp + 255) / 256);
    hipLaunchKernelGGL(colmax_kernel, dim3(bx, nslices), dim3(256), 0, s, X, ld, R, K, kslice, part, Rp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(colscale_kernel, dim3(bx), dim3(256), 0, s, part, nslices, Rp, scale, inv);
    return hipGetLastError();
}

struct Layout {
    int64_t Mp, Np, Kp;
    int np;
    uint16_t *Ap, *Bp;
    float *a_scale, *a_inv, *b_scale, *b_inv;
    float *part;  // column-maximum partials, kScaleSlices x max(Mp, Np)
};

Layout layout(int64_t M, int64_t N, int64_t K, int np, bool scaled, void *work)
{
    Layout L;
    L.Mp = round_up(M, kRowPad);
    L.Np = round_up(N, kRowPad);
    L.Kp = round_up(K, kKpad);
    L.np = np;
    L.Ap = (uint16_t *)work;
    L.Bp = L.Ap + (size_t)np * L.Mp * L.Kp;
    float *f = (float *)(L.Bp + (size_t)np * L.Np * L.Kp);
    L.a_scale = scaled ? f : nullptr;
    L.a_inv = scaled ? f + L.Mp : nullptr;
    L.b_scale = scaled ? f + 2 * L.Mp : nullptr;
    L.b_inv = scaled ? f + 2 * L.Mp + L.Np : nullptr;
    L.part = scaled ? f + 2 * (L.Mp + L.Np) : nullptr;
    return L;
}

template <int S>
hipError_t run_scheme(const Gemm<float> &g, void *work, hipStream_t s)
{
    using Sc = Scheme<S>;
    const Layout L = layout(g.M, g.N, g.K, Sc::NP, Sc::FP16, work);
    hipError_t e = hipSuccess;
    // op(A) rows are M x K: k contiguous unless A is transposed.
    // op(B)^T rows are N x K: k contiguous when B is transposed.
    if (Sc::FP16) {
        e = launch_rowscale(g.A, g.lda, !g.ta, g.M, g.K, L.a_scale, L.a_inv, L.Mp, L.part, s);
        if (e == hipSuccess) e = launch_rowscale(g.B, g.ldb, g.tb, g.N, g.K, L.b_scale, L.b_inv, L.Np, L.part, s);
    }
    if (e == hipSuccess) e = launch_split<Sc::NP, Sc::FP16>(g.A, g.lda, !g.ta, g.M, g.K, L.Ap, L.Mp, L.Kp, L.a_scale, s);
    if (e == hipSuccess) e = launch_split<Sc::NP, Sc::FP16>(g.B, g.ldb, g.tb, g.N, g.K, L.Bp, L.Np, L.Kp, L.b_scale, s);
    if (e != hipSuccess) return e;
    const int64_t tiles_m = L.Mp / kBM, tiles_n = L.Np / kBN;
    const int64_t nwg = tiles_m * tiles_n;
    if (nwg > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(planes_kernel<S>, dim3((unsigned)nwg), dim3(256), 0, s, L.Ap, L.Bp, L.Kp, L.Mp * L.Kp,
                       L.Np * L.Kp, (int)tiles_m, (int)tiles_n, g.C, g.ldc, g.M, g.N, g.alpha, g.beta, L.a_inv,
                       L.b_inv);
    return hipGetLastError();
}

}  // namespace

size_t split_workspace_bytes(int64_t M, int64_t N, int64_t K, int scheme)
{
    const int np = scheme == kBF16x6 ? 3 : 2;
    const bool scaled = scheme == kFP16x3;
    // the six-product form's kernel takes op(A) in 256-row tiles
    const int64_t Mp = round_up(M, scheme == kBF16x6 ? 2 * kRowPad : kRowPad), Np = round_up(N, kRowPad),
                  Kp = round_up(K, kKpad);
    return (size_t)np * (size_t)(Mp + Np) * (size_t)Kp * sizeof(uint16_t) +
           (scaled ? ((size_t)2 * (Mp + Np) + (size_t)kScaleSlices * (Mp > Np ? Mp : Np)) * sizeof(float) : 0);
}

hipError_t launch_split_gemm(const Gemm<float> &g, int scheme, void *work, hipStream_t s)
{
    switch (scheme) {
    case kBF16x6: return launch_split6_frag(g, work, s);
    case kBF16x3: return run_scheme<kBF16x3>(g, work, s);
    case kFP16x3: return run_scheme<kFP16x3>(g, work, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace blasdsp
