// gemm_split.hip — FP32 GEMM on the BF16 matrix cores ("split-BF16").
//
// No counterpart in the reference (an added precision form).  Every FP32
// operand x is split into BF16 pieces x = hi + mid + lo (+ a residual below
// 2^-24 |x|), each piece the round-to-nearest-even BF16 of what the previous
// pieces left over.  The product is assembled from the piece products whose
// magnitude is at least 2^-16 of the full product:
//   6 terms:  hh + hm + mh + hl + lh + mm   (three pieces; residual ~2^-24)
//   3 terms:  hh + hl + lh with two pieces hi + lo (residual ~2^-17)
// All products are exact in the FP32 MFMA accumulator inputs (8-bit x 8-bit
// significands), so the error comes from the dropped terms, the split residual
// and FP32 accumulation.  Accumulation is two-level: each 256-deep K chunk is
// accumulated from zero in its own MFMA accumulators and then added into the
// running sum, which keeps the rounding error of a K = 8192 sum near that of a
// 32-term sum instead of an 8192-term one.
//
// Two kernels:
//   split_kernel   reads op(A) (or op(B)^T) in any storage order and writes the
//                  pieces as BF16 planes [rows][Kp], K contiguous, zero padded
//                  to whole 128-row tiles and 64-deep K blocks — the transpose
//                  modes are absorbed here, through an LDS tile;
//   planes_kernel  a 128 x 128-tile MFMA GEMM (v_mfma_f32_16x16x32_bf16) over
//                  the planes: per 32-deep K step each wave reads every piece
//                  fragment once from LDS and issues nterms x 16 MFMAs, so the
//                  LDS traffic per MFMA falls as the number of terms grows.
#include "internal.h"

namespace blasdsp {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kBM = 128, kBN = 128, kBK = 32;  // GEMM tile
constexpr int kKpad = 64;                      // plane K padding (split tile)
constexpr int kRowPad = 128;                   // plane row padding
constexpr int kLdsStride = kBK + 8;            // bf16 per LDS row (80 bytes)
constexpr int kChunkTiles = 8;                 // K tiles per accumulation chunk (256)

// Round-to-nearest-even FP32 -> BF16 bits (finite inputs).
__device__ __forceinline__ uint32_t bf16_bits(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ float bf16_val(uint32_t h) { return __uint_as_float(h << 16); }

// Split x into npieces BF16 pieces (bits).
template <int NP>
__device__ __forceinline__ void split(float x, uint32_t (&p)[NP])
{
    float r = x;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        p[i] = bf16_bits(r);
        r = r - bf16_val(p[i]);  // exact: r and its BF16 rounding agree in the top bits
    }
}

// One 64 x 64 tile of op(X) (R logical rows x K) into NP planes [Rp][Kp].
// KC: X is stored [R][K] (k contiguous); else [K][R].
template <int NP, bool KC>
__global__ void __launch_bounds__(256) split_kernel(const float *__restrict__ X, int64_t ld, int64_t R,
                                                   int64_t K, uint16_t *__restrict__ planes,
                                                   int64_t Kp, int64_t plane_elems)
{
    __shared__ float tile[64][65];
    const int tid = threadIdx.x;
    const int64_t k0 = (int64_t)blockIdx.x * 64, r0 = (int64_t)blockIdx.y * 64;
    if (KC) {
        // direct: thread covers (row, 4 consecutive k) pieces
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int e = tid + i * 256;  // 1024 groups of 4
            const int row = e / 16, kq = (e % 16) * 4;
            const int64_t gr = r0 + row;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int64_t gk = k0 + kq + j;
                const float x = (gr < R && gk < K) ? X[gr * ld + gk] : 0.0f;
                uint32_t p[NP];
                split<NP>(x, p);
#pragma unroll
                for (int q = 0; q < NP; ++q) planes[q * plane_elems + gr * Kp + gk] = (uint16_t)p[q];
            }
        }
    } else {
        // transposed: coalesced reads along r into LDS, then writes along k
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int e = tid + i * 256;
            const int kk = e / 64, rr = e % 64;
            const int64_t gk = k0 + kk, gr = r0 + rr;
            tile[kk][rr] = (gr < R && gk < K) ? X[gk * ld + gr] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int e = tid + i * 256;
            const int rr = e / 64, kk = e % 64;
            const float x = tile[kk][rr];
            uint32_t p[NP];
            split<NP>(x, p);
            const int64_t gr = r0 + rr, gk = k0 + kk;
#pragma unroll
            for (int q = 0; q < NP; ++q) planes[q * plane_elems + gr * Kp + gk] = (uint16_t)p[q];
        }
    }
}

// Product list: pairs (piece of A, piece of B), small terms first.
template <int NTERMS>
struct Terms;
template <>
struct Terms<6> {
    static constexpr int NP = 3;
    static constexpr int a[6] = {1, 0, 2, 0, 1, 0};
    static constexpr int b[6] = {1, 2, 0, 1, 0, 0};
};
template <>
struct Terms<3> {
    static constexpr int NP = 2;
    static constexpr int a[3] = {0, 1, 0};
    static constexpr int b[3] = {1, 0, 0};
};

// f32 16x16 output of the bf16 MFMA: lane group g, register r hold row 4 g + r.
__device__ __forceinline__ int out_row(int g, int r) { return 4 * g + r; }

template <int NTERMS>
__global__ void __launch_bounds__(256) planes_kernel(const uint16_t *__restrict__ Ap,
                                                     const uint16_t *__restrict__ Bp, int64_t Kp,
                                                     int64_t a_plane, int64_t b_plane, int tiles_m,
                                                     int tiles_n, float *__restrict__ C, int64_t ldc,
                                                     int64_t M, int64_t N, float alpha, float beta)
{
    using Tm = Terms<NTERMS>;
    constexpr int NP = Tm::NP;
    constexpr int PLANE_LDS = kBM * kLdsStride;           // bf16 per plane per stage
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

    // staging: per plane and operand 128 rows x 32 bf16 = 512 16-byte vectors, 2 per thread
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

    f4 acc[4][4], part[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = part[i][j] = f4{0, 0, 0, 0};

    const int nk = (int)(Kp / kBK);
    gload(0);
    lstore(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) gload((int64_t)(kt + 1) * kBK);

        bf16x8 fa[NP][4], fb[NP][4];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                fa[p][i] = *reinterpret_cast<const bf16x8 *>(
                    &lds[cur][0][p][(wm + i * 16 + lr) * kLdsStride + lg * 8]);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                fb[p][j] = *reinterpret_cast<const bf16x8 *>(
                    &lds[cur][1][p][(wn + j * 16 + lr) * kLdsStride + lg * 8]);
        }
#pragma unroll
        for (int t = 0; t < NTERMS; ++t)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    part[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[Tm::a[t]][i], fb[Tm::b[t]][j],
                                                                        part[i][j], 0, 0, 0);
        if ((kt + 1) % kChunkTiles == 0 || !more) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    acc[i][j] += part[i][j];
                    part[i][j] = f4{0, 0, 0, 0};
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
                    float *p = C + row * ldc + col;
                    float v = alpha * acc[i][j][r];
                    if (beta != 0.0f) v += beta * *p;
                    *p = v;
                }
            }
        }
}

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

template <int NP>
hipError_t launch_split(const float *X, int64_t ld, bool kc, int64_t R, int64_t K, uint16_t *planes,
                        int64_t Rp, int64_t Kp, hipStream_t s)
{
    const dim3 grid((unsigned)(Kp / 64), (unsigned)(Rp / 64));
    if (kc)
        hipLaunchKernelGGL((split_kernel<NP, true>), grid, dim3(256), 0, s, X, ld, R, K, planes, Kp, Rp * Kp);
    else
        hipLaunchKernelGGL((split_kernel<NP, false>), grid, dim3(256), 0, s, X, ld, R, K, planes, Kp, Rp * Kp);
    return hipGetLastError();
}

}  // namespace

size_t split_bf16_workspace_bytes(int64_t M, int64_t N, int64_t K, int nterms)
{
    const int np = nterms == 6 ? 3 : 2;
    const int64_t Mp = round_up(M, kRowPad), Np = round_up(N, kRowPad), Kp = round_up(K, kKpad);
    return (size_t)np * (size_t)(Mp + Np) * (size_t)Kp * sizeof(uint16_t);
}

hipError_t launch_split_bf16(const Gemm<float> &g, int nterms, void *work, hipStream_t s)
{
    if (nterms != 3 && nterms != 6) return hipErrorInvalidValue;
    const int64_t Mp = round_up(g.M, kRowPad), Np = round_up(g.N, kRowPad), Kp = round_up(g.K, kKpad);
    const int np = nterms == 6 ? 3 : 2;
    uint16_t *Ap = (uint16_t *)work;
    uint16_t *Bp = Ap + (size_t)np * Mp * Kp;
    hipError_t e;
    // op(A) rows are M x K: k contiguous unless A is transposed.
    // op(B)^T rows are N x K: k contiguous when B is transposed.
    if (np == 3) {
        e = launch_split<3>(g.A, g.lda, !g.ta, g.M, g.K, Ap, Mp, Kp, s);
        if (e == hipSuccess) e = launch_split<3>(g.B, g.ldb, g.tb, g.N, g.K, Bp, Np, Kp, s);
    } else {
        e = launch_split<2>(g.A, g.lda, !g.ta, g.M, g.K, Ap, Mp, Kp, s);
        if (e == hipSuccess) e = launch_split<2>(g.B, g.ldb, g.tb, g.N, g.K, Bp, Np, Kp, s);
    }
    if (e != hipSuccess) return e;
    const int64_t tiles_m = Mp / kBM, tiles_n = Np / kBN;
    const int64_t nwg = tiles_m * tiles_n;
    if (nwg > 0x7fffffffLL) return hipErrorInvalidValue;
    if (nterms == 6)
        hipLaunchKernelGGL(planes_kernel<6>, dim3((unsigned)nwg), dim3(256), 0, s, Ap, Bp, Kp, Mp * Kp,
                           Np * Kp, (int)tiles_m, (int)tiles_n, g.C, g.ldc, g.M, g.N, g.alpha, g.beta);
    else
        hipLaunchKernelGGL(planes_kernel<3>, dim3((unsigned)nwg), dim3(256), 0, s, Ap, Bp, Kp, Mp * Kp,
                           Np * Kp, (int)tiles_m, (int)tiles_n, g.C, g.ldc, g.M, g.N, g.alpha, g.beta);
    return hipGetLastError();
}

}  // namespace blasdsp
