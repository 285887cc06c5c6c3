// gemm_split.hip — FP32 GEMM on the 16-bit matrix cores ("split" forms):
// the forms, their scales and workspace; the GEMM itself is in
// gemm_split_frag.hip.
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
// BF16 forms accumulate in two levels: each 192-deep K chunk from zero in its
// own MFMA accumulators, then added into the running sum.
//
// Kernels here: rowscale_rows_kernel / colmax_kernel + colscale_kernel (FP16
// form) compute the per-row power-of-two scales of op(A) and op(B)^T.
#include "internal.h"

namespace blasdsp {
namespace {

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

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

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

enum { kBF16x6 = 0, kBF16x3 = 1, kFP16x3 = 2 };
constexpr int kKpad = 64;  // K padding of the pieces

// Workspace: op(A)'s pieces (np x Mp x Kp), op(B)^T's (np x Np x Kp), then
// for the FP16 form the scales of op(A) and op(B)^T, their inverses and the
// column-maximum partials (kScaleSlices x max(Mp, Np)).  Rows of op(A) are
// padded to 256, of op(B)^T to 128, K to 64 (the GEMM's 256 x 128 tiles and
// the split's 64-deep blocks).
struct Layout {
    int64_t Mp, Np, Kp;
    int np;
    uint16_t *Ap, *Bp;
    float *a_scale, *a_inv, *b_scale, *b_inv;
    float *part;
};

Layout layout(int64_t M, int64_t N, int64_t K, int np, bool scaled, void *work)
{
    Layout L;
    L.Mp = round_up(M, 256);
    L.Np = round_up(N, 128);
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

}  // namespace

size_t split_workspace_bytes(int64_t M, int64_t N, int64_t K, int scheme)
{
    const int np = scheme == kBF16x6 ? 3 : 2;
    const bool scaled = scheme == kFP16x3;
    const int64_t Mp = round_up(M, 256), Np = round_up(N, 128), Kp = round_up(K, kKpad);
    return (size_t)np * (size_t)(Mp + Np) * (size_t)Kp * sizeof(uint16_t) +
           (scaled ? ((size_t)2 * (Mp + Np) + (size_t)kScaleSlices * (Mp > Np ? Mp : Np)) * sizeof(float) : 0);
}

hipError_t launch_split_gemm(const Gemm<float> &g, int scheme, void *work, hipStream_t s)
{
    if (scheme != kBF16x6 && scheme != kBF16x3 && scheme != kFP16x3) return hipErrorInvalidValue;
    const bool scaled = scheme == kFP16x3;
    const Layout L = layout(g.M, g.N, g.K, scheme == kBF16x6 ? 3 : 2, scaled, work);
    hipError_t e = hipSuccess;
    // op(A) rows are M x K: k contiguous unless A is transposed.
    // op(B)^T rows are N x K: k contiguous when B is transposed.
    if (scaled) {
        e = launch_rowscale(g.A, g.lda, !g.ta, g.M, g.K, L.a_scale, L.a_inv, L.Mp, L.part, s);
        if (e == hipSuccess) e = launch_rowscale(g.B, g.ldb, g.tb, g.N, g.K, L.b_scale, L.b_inv, L.Np, L.part, s);
        if (e != hipSuccess) return e;
    }
    return launch_frag_gemm(g, scheme, L.Ap, L.Bp, L.Mp, L.Np, L.Kp, L.a_scale, L.a_inv, L.b_scale, L.b_inv, s);
}

}  // namespace blasdsp
