// internal.h — shared declarations between the C-ABI layer and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace blasdsp {

// One GEMM, already validated: C = alpha * op(A) op(B) + beta * C (row-major).
template <typename T>
struct Gemm {
    bool ta, tb;          // op(A) = A^T, op(B) = B^T
    int64_t M, N, K;
    T alpha, beta;
    const T *A;
    int64_t lda;
    const T *B;
    int64_t ldb;
    T *C;
    int64_t ldc;
};

// MFMA tile kernel (gemm_tile.hip): any shape, any mode; bm = 64 or 128 is the
// workgroup tile edge.
// MFMA tile kernel with bm x bm tiles (64 or 128).  splitk > 1 cuts K into
// that many slices (whole K tiles); their partials go to `work` (splitk * M *
// N elements of T) and a second kernel adds them in order and applies the
// epilogue.
template <typename T>
hipError_t launch_tile(const Gemm<T> &g, int bm, int splitk, T *work, hipStream_t s);
// K slices the tile kernel would use for this shape on this GPU (1: none).
template <typename T>
int tile_pick_splitk(const Gemm<T> &g, int bm, int num_cus);

// C = beta * C over M x N (beta == 0 writes zeros): K == 0 or alpha == 0.
template <typename T>
hipError_t launch_scale(const Gemm<T> &g, hipStream_t s);

// Empty kernel, for launch-overhead measurement.
hipError_t launch_null(hipStream_t s);

// Skinny (shortA / thinB) kernels (gemm_skinny.hip).  `splitk` >= 1; when > 1
// the partial sums go to `work` (splitk * M * N elements of T).
template <typename T>
hipError_t launch_skinny(const Gemm<T> &g, int splitk, T *work, hipStream_t s);
// Whether the skinny kernel takes this shape: small side <= 64, and strides
// small enough that its per-lane load offsets stay below 2^31 bytes.
template <typename T>
bool skinny_fits(const Gemm<T> &g);
// Number of K slices the skinny kernel would use for this shape on this GPU.
template <typename T>
int skinny_pick_splitk(const Gemm<T> &g, int num_cus);
// Whether the skinny kernel's panels and K slices fill one round of the
// GPU's workgroup slots (else the tile kernel is the faster choice).
template <typename T>
bool skinny_fills(const Gemm<T> &g, int num_cus);

// FP32 GEMM on 16-bit matrix cores (gemm_split.hip).  scheme: 0 = BF16 pieces
// hi/mid/lo with 6 products, 1 = BF16 hi/lo with 3 products, 2 = scaled FP16
// hi/lo with 3 products.  `work` must hold split_workspace_bytes(...) bytes.
size_t split_workspace_bytes(int64_t M, int64_t N, int64_t K, int scheme);
hipError_t launch_split_gemm(const Gemm<float> &g, int scheme, void *work, hipStream_t s);
// The split forms on fragment-tiled pieces (gemm_split_frag.hip): splits
// op(A) into Ap (pieces x Mp x Kp) and op(B)^T into Bp (pieces x Np x Kp),
// then runs the GEMM.  Mp a multiple of 256, Np of 128, Kp of 64.  The scaled
// FP16 form (scheme 2) takes power-of-two row scales of op(A) and op(B)^T and
// their inverses (Mp and Np entries).
hipError_t launch_frag_gemm(const Gemm<float> &g, int scheme, uint16_t *Ap, uint16_t *Bp, int64_t Mp, int64_t Np,
                            int64_t Kp, const float *a_scale, const float *a_inv, const float *b_scale,
                            const float *b_inv, hipStream_t s);

// FP32 GEMM on the FP8 (E4M3) matrix cores with per-128-block power-of-two
// scales (gemm_fp8.hip).  `work` must hold fp8_workspace_bytes(...) bytes.
size_t fp8_workspace_bytes(int64_t M, int64_t N, int64_t K);
hipError_t launch_fp8_gemm(const Gemm<float> &g, void *work, hipStream_t s);

// Synthetic input generator (fill.hip).
template <typename T>
hipError_t launch_fill(T *X, int64_t rows, int64_t cols, int64_t ld, int trans, int dist,
                       uint64_t seed, T pad, hipStream_t s);

}  // namespace blasdsp
