// api_common.h — helpers shared by the C-ABI translation units.
#pragma once

#include "../../include/blasdsp.h"
#include "internal.h"

namespace blasdsp {

extern thread_local int g_last_hip_error;

// hipError_t -> BLASDSP_* status (records the HIP code for blasdsp_last_hip_error).
int hip_status(hipError_t e);
// Compute units of the current device (cached).
int num_cus();
// Grow-only scratch buffer owned by the library, one per (device, stream).
hipError_t workspace(hipStream_t s, size_t bytes, void **out);
// Frees the current device's scratch buffer of stream s (if any); the next
// workspace() call for s allocates afresh.  The caller guarantees no work
// using it is pending (hipFree also waits for the device).
hipError_t workspace_release(hipStream_t s);
// Destroys the cached group plan of stream s on the current device (if any).
hipError_t grouped_plan_release(hipStream_t s);
bool parse_trans(char t, bool &out);

// Kernel and split-K choice of one call (AUTO dispatch, tuned table, or the
// caller's explicit algo / splitk), shared by every entry point that runs the
// tile or skinny kernel.
struct KernelChoice {
    bool skinny;   // skinny kernel (else the tile kernel)
    int tile;      // tile edge of the tile kernel (64 or 128)
    int splitk;    // K slices (1: none)
};
// Validated g with M, N > 0, K > 0 and alpha != 0.  splitk <= 0: the library
// picks the slice count.
template <typename T>
int choose_kernel(const Gemm<T> &g, int algo, int splitk, KernelChoice &c);
// Bytes of split-K partials a call with `splitk` slices may write (an upper
// bound: the kernels round slices to whole K blocks, which can only lower
// the count).
template <typename T>
size_t kernel_scratch_bytes(const Gemm<T> &g, int splitk);
// One GEMM through the library's dispatch (argument checks, kernel choice,
// workspace): what blasdsp_dgemm_ex / blasdsp_sgemm_ex run.
template <typename T>
int gemm_entry(char ta, char tb, int64_t M, int64_t N, int64_t K, T alpha, const T *A, int64_t lda,
               const T *B, int64_t ldb, T beta, T *C, int64_t ldc, int algo, int splitk, void *stream);
// Checks the arguments as gemm_entry does, then launches the empty kernel.
template <typename T>
int null_entry(char ta, char tb, int64_t lda, int64_t ldb, int64_t ldc, int64_t M, int64_t N,
               int64_t K, const T *A, const T *B, T *C, T alpha, T beta, void *stream);
// Enqueues the chosen kernel; `work` holds kernel_scratch_bytes(g, c.splitk).
template <typename T>
int launch_kernel(const Gemm<T> &g, const KernelChoice &c, T *work, hipStream_t s);

// Argument checks shared by every entry point (reference shape limits,
// dgemm_release_V1.1/readme.txt:11-15, are not carried over: any M, N, K >= 0
// and any leading dimension covering a row are accepted).
template <typename T>
int check(Gemm<T> &g, char transA, char transB, int64_t M, int64_t N, int64_t K, T alpha,
          const T *A, int64_t lda, const T *B, int64_t ldb, T beta, T *C, int64_t ldc)
{
    if (!parse_trans(transA, g.ta) || !parse_trans(transB, g.tb)) return BLASDSP_ERR_INVALID;
    if (M < 0 || N < 0 || K < 0) return BLASDSP_ERR_INVALID;
    const int64_t a_cols = g.ta ? M : K, b_cols = g.tb ? K : N;
    if (lda < (a_cols > 1 ? a_cols : 1)) return BLASDSP_ERR_INVALID;
    if (ldb < (b_cols > 1 ? b_cols : 1)) return BLASDSP_ERR_INVALID;
    if (ldc < (N > 1 ? N : 1)) return BLASDSP_ERR_INVALID;
    if (M > 0 && N > 0) {
        if (!C) return BLASDSP_ERR_INVALID;
        if (K > 0 && alpha != T(0) && (!A || !B)) return BLASDSP_ERR_INVALID;
    }
    g.M = M; g.N = N; g.K = K;
    g.alpha = alpha; g.beta = beta;
    g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
    return BLASDSP_SUCCESS;
}


}  // namespace blasdsp
