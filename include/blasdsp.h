/*
 * blasdsp.h — C ABI of the MI355X (gfx950) GEMM library.
 *
 * Every GEMM computes, row-major,
 *
 *     C = alpha * op(A) * op(B) + beta * C
 *
 * op(A) is M x K and op(B) is K x N.  For transA == 'N' A is stored M x K
 * (lda >= K); for 'T' A is stored K x M (lda >= M).  Likewise B is K x N
 * (ldb >= N) for 'N' and N x K (ldb >= K) for 'T'; C is M x N (ldc >= N).
 * Element (i, j) of a matrix with leading dimension ld is X[i * ld + j], as in
 * the reference (#define OFFSET(row, col, ld), dgemm_release_V1.1/device_code/
 * dgemm_kernel.c:8).  When beta == 0, C is not read.
 *
 * A, B and C are device (HBM) pointers owned by the caller.  Every call is
 * asynchronous on `stream` (a hipStream_t; NULL = the null stream) and
 * returns a status code; errors detected on the host leave C untouched.
 * Calls on distinct streams may run concurrently.
 *
 * Semantics relative to the reference.  The reference's dgemm and irr kernels
 * compute C += op(A)op(B); sgemm V1.3.2 and irr_thinB_sgemm V1.2 compute
 * C = alpha*op(A)op(B) + beta with a SCALAR beta and ignore the input C.  The
 * entry points that carry the reference's packed-operand names and argument
 * lists (sgemm_*_qxx, irr_*) keep those semantics, so they are drop-ins;
 * dgemm_NN..TT keep the reference's leading-dimension argument list with
 * alpha and beta appended (1, 1 reproduce C += op(A)op(B)); every other entry
 * point implements the BLAS form above.
 */
#ifndef BLASDSP_H
#define BLASDSP_H

#include <stdint.h>

#if defined(__GNUC__)
#define BLASDSP_API __attribute__((visibility("default")))
#else
#define BLASDSP_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define BLASDSP_SUCCESS 0
#define BLASDSP_ERR_INVALID (-22)     /* bad dimension, leading dim, mode or NULL */
#define BLASDSP_ERR_NOMEM (-12)       /* workspace / budget allocation failed */
#define BLASDSP_ERR_DEVICE (-5)       /* HIP runtime error (launch, copy, sync) */
#define BLASDSP_ERR_UNSUPPORTED (-95) /* valid request the library does not serve */

BLASDSP_API const char *blasdsp_version(void);
BLASDSP_API const char *blasdsp_status_string(int status);
/* Last HIP error code seen by the library on this thread (0 if none). */
BLASDSP_API int blasdsp_last_hip_error(void);

/* ---- general entry points ----------------------------------------------- */

/* Dispatches by shape: the MFMA tile kernel for general shapes, the
 * bandwidth-bound skinny kernel when M or N is at most 48 and the other side
 * is long enough for the skinny kernel's panels and K slices to fill the GPU;
 * a shape class recorded by blasdsp_autotune takes its recorded choice.
 * Replaces the per-mode __global__ functions of
 * dgemm_release_V1.1/include/dgemm_kernel.h:11-38. */
BLASDSP_API int blasdsp_dgemm(char transA, char transB, int64_t M, int64_t N, int64_t K,
                  double alpha, const double *A, int64_t lda,
                  const double *B, int64_t ldb, double beta,
                  double *C, int64_t ldc, void *stream);

/* FP32, FP32 accumulation on MFMA.  Replaces sgemm_NN_qxx ... sgemm_TT_qxx,
 * sgemm_release_V1.3.2/include/sgemm_kernel.h:11-23. */
BLASDSP_API int blasdsp_sgemm(char transA, char transB, int64_t M, int64_t N, int64_t K,
                  float alpha, const float *A, int64_t lda,
                  const float *B, int64_t ldb, float beta,
                  float *C, int64_t ldc, void *stream);

/* ---- dgemm V1.1 per-mode entry points (reference names) ------------------
 * Argument order follows the reference (lda, ldb, ldc, M, N, K, A, B, C) with
 * the barrier ids dropped (no group barriers exist on the GPU) and alpha,
 * beta, stream appended (alpha/beta last, as in sgemm_NN_qxx).
 *   reference dgemm_NN(lda, ldb, ldc, M, N, K, b_id, A, B, C)   C += A B
 *   ->        dgemm_NN(lda, ldb, ldc, M, N, K, A, B, C, 1.0, 1.0, stream)   */

/* dgemm_release_V1.1/include/dgemm_kernel.h:11 */
BLASDSP_API int dgemm_NN(int64_t lda, int64_t ldb, int64_t ldc, int64_t M, int64_t N, int64_t K,
             const double *A, const double *B, double *C, double alpha, double beta,
             void *stream);
/* dgemm_release_V1.1/include/dgemm_kernel.h:18 */
BLASDSP_API int dgemm_NT(int64_t lda, int64_t ldb, int64_t ldc, int64_t M, int64_t N, int64_t K,
             const double *A, const double *B, double *C, double alpha, double beta,
             void *stream);
/* dgemm_release_V1.1/include/dgemm_kernel.h:25 (bid_main, bid_T dropped) */
BLASDSP_API int dgemm_TN(int64_t lda, int64_t ldb, int64_t ldc, int64_t M, int64_t N, int64_t K,
             const double *A, const double *B, double *C, double alpha, double beta,
             void *stream);
/* dgemm_release_V1.1/include/dgemm_kernel.h:33 */
BLASDSP_API int dgemm_TT(int64_t lda, int64_t ldb, int64_t ldc, int64_t M, int64_t N, int64_t K,
             const double *A, const double *B, double *C, double alpha, double beta,
             void *stream);
/* dgemm_release_V1.1/include/dgemm_kernel.h:40 — empty kernel launch, used to
 * measure launch overhead as the reference harness does
 * (dgemm_release_V1.1/host_code/dgemm_demo.c:140-146).  Arguments are checked
 * like dgemm_NN's and otherwise ignored. */
BLASDSP_API int dgemm_null(int64_t lda, int64_t ldb, int64_t ldc, int64_t M, int64_t N, int64_t K,
               const double *A, const double *B, double *C, double alpha, double beta,
               void *stream);

/* ---- the reference's packed-operand kernels ------------------------------
 * The reference's names, argument order and semantics, with the barrier ids
 * (b_id; bid_main, bid_T) dropped and a stream appended.  Operands are
 * packed as in the reference's demos: A is M x K (lda = K) or, for the T
 * modes, K x M (lda = M); B is K x N (ldb = N) or N x K (ldb = K); C is
 * M x N (ldc = N).  The ld-taking BLAS-semantics generalisations are the
 * blasdsp_* entries below.
 *
 * sgemm V1.3.2: C = alpha op(A) op(B) + beta, beta a SCALAR added to every
 * element, the input C ignored (sgemm_release_V1.3.2/device_code/
 * sgemm_kernel.c:2681,2745-2749).  Implemented as C = beta (one write pass
 * over C, the reference's C_plus_beta, sgemm_kernel.c:25-160), then the BLAS
 * GEMM with beta = 1. */
/* sgemm_release_V1.3.2/include/sgemm_kernel.h:11 */
BLASDSP_API int sgemm_NN_qxx(int64_t M, int64_t N, int64_t K, const float *A, const float *B, float *C,
                             float alpha, float beta, void *stream);
/* sgemm_release_V1.3.2/include/sgemm_kernel.h:15 */
BLASDSP_API int sgemm_NT_qxx(int64_t M, int64_t N, int64_t K, const float *A, const float *B, float *C,
                             float alpha, float beta, void *stream);
/* sgemm_release_V1.3.2/include/sgemm_kernel.h:18 (bid_main, bid_T dropped) */
BLASDSP_API int sgemm_TN_qxx(int64_t M, int64_t N, int64_t K, const float *A, const float *B, float *C,
                             float alpha, float beta, void *stream);
/* sgemm_release_V1.3.2/include/sgemm_kernel.h:22 */
BLASDSP_API int sgemm_TT_qxx(int64_t M, int64_t N, int64_t K, const float *A, const float *B, float *C,
                             float alpha, float beta, void *stream);
/* sgemm_release_V1.3.2/include/sgemm_kernel.h:25 — empty kernel launch;
 * arguments are checked like sgemm_NN_qxx's and otherwise ignored. */
BLASDSP_API int sgemm_null(int64_t M, int64_t N, int64_t K, const float *A, const float *B, float *C,
                           float alpha, float beta, void *stream);

/* irr_shortA (small M) and irr_thinB (small N): C += op(A) op(B), except
 * irr_thinB_sgemm V1.2, which computes C = alpha op(A) op(B) + beta (scalar
 * beta, input C ignored; irr_thinB_sgemm_release_V1.2/device_code/
 * irr_thinB_sgemm_kernel.c:3095-3112,3182-3193), done as for sgemm_*_qxx.
 * They run the bandwidth-bound skinny kernel (op(B) or op(A) streamed once
 * from HBM, the small operand held on chip, split-K with a deterministic
 * reduction when the long side alone does not fill the GPU); shapes it does
 * not take (small side above 64, or per-lane load offsets at or past 2^31
 * bytes) run on the tile kernel.  The *_null entries launch the empty kernel
 * after checking their arguments. */
/* irr_shortA_dgemm_release_V0.0/include/irr_shortA_dgemm_kernel.h:11,15,18,22,25 */
BLASDSP_API int irr_shortA_dgemm_NN_dbuf(int64_t M, int64_t N, int64_t K, const double *A, const double *B,
                                         double *C, void *stream);
BLASDSP_API int irr_shortA_dgemm_NT_dbuf(int64_t M, int64_t N, int64_t K, const double *A, const double *B,
                                         double *C, void *stream);
BLASDSP_API int irr_shortA_dgemm_TN_dbuf(int64_t M, int64_t N, int64_t K, const double *A, const double *B,
                                         double *C, void *stream);
BLASDSP_API int irr_shortA_dgemm_TT_dbuf(int64_t M, int64_t N, int64_t K, const double *A, const double *B,
                                         double *C, void *stream);
BLASDSP_API int irr_shortA_dgemm_null(int64_t M, int64_t N, int64_t K, const double *A, const double *B,
                                      double *C, void *stream);
/* irr_shortA_sgemm_release_V0.0/include/irr_shortA_sgemm_kernel.h:11,15,18,22,25 */
BLASDSP_API int irr_shortA_sgemm_NN(int64_t M, int64_t N, int64_t K, const float *A, const float *B, float *C,
                                    void *stream);
BLASDSP_API int irr_shortA_sgemm_NT(int64_t M, int64_t N, int64_t K, const float *A, const float *B, float *C,
                                    void *stream);
BLASDSP_API int irr_shortA_sgemm_TN(int64_t M, int64_t N, int64_t K, const float *A, const float *B, float *C,
                                    void *stream);
BLASDSP_API int irr_shortA_sgemm_TT(int64_t M, int64_t N, int64_t K, const float *A, const float *B, float *C,
                                    void *stream);
BLASDSP_API int irr_shortA_sgemm_null(int64_t M, int64_t N, int64_t K, const float *A, const float *B, float *C,
                                      void *stream);
/* irr_thinB_dgemm_release_V0.0/include/irr_thinB_dgemm_kernel.h:11-25 (the
 * header's comments on the TT and NN prototypes are swapped; the names are
 * what the kernels compute) */
BLASDSP_API int irr_thinB_dgemm_NN_dbuf(int64_t M, int64_t N, int64_t K, const double *A, const double *B,
                                        double *C, void *stream);
BLASDSP_API int irr_thinB_dgemm_NT_dbuf(int64_t M, int64_t N, int64_t K, const double *A, const double *B,
                                        double *C, void *stream);
BLASDSP_API int irr_thinB_dgemm_TN_dbuf(int64_t M, int64_t N, int64_t K, const double *A, const double *B,
                                        double *C, void *stream);
BLASDSP_API int irr_thinB_dgemm_TT_dbuf(int64_t M, int64_t N, int64_t K, const double *A, const double *B,
                                        double *C, void *stream);
BLASDSP_API int irr_thinB_dgemm_null(int64_t M, int64_t N, int64_t K, const double *A, const double *B,
                                     double *C, void *stream);
/* irr_thinB_sgemm_release_V1.2/include/irr_thinB_sgemm_kernel.h:11-25 */
BLASDSP_API int irr_thinB_sgemm_NN(int64_t M, int64_t N, int64_t K, const float *A, const float *B, float *C,
                                   float alpha, float beta, void *stream);
BLASDSP_API int irr_thinB_sgemm_NT(int64_t M, int64_t N, int64_t K, const float *A, const float *B, float *C,
                                   float alpha, float beta, void *stream);
BLASDSP_API int irr_thinB_sgemm_TN(int64_t M, int64_t N, int64_t K, const float *A, const float *B, float *C,
                                   float alpha, float beta, void *stream);
BLASDSP_API int irr_thinB_sgemm_TT(int64_t M, int64_t N, int64_t K, const float *A, const float *B, float *C,
                                   float alpha, float beta, void *stream);
BLASDSP_API int irr_thinB_sgemm_null(int64_t M, int64_t N, int64_t K, const float *A, const float *B, float *C,
                                     float alpha, float beta, void *stream);

/* ---- by-name launch -------------------------------------------------------
 * Replaces the reference host's hthread_group_exec(tg_id, name, n_scalar,
 * n_ptr, args) (dgemm_release_V1.1/host_code/dgemm_demo.c:151): runs the
 * launcher `name` (launch_dgemm_NN ... launch_irr_thinB_sgemm_null, the names
 * of the reference's *_launcher.c files) on the same argument words —
 * n_scalar 64-bit scalars (dimensions, then barrier ids, which are ignored)
 * followed by n_ptr = 3 device pointers A, B, C.  The counts must be the
 * reference launcher's (7 for launch_dgemm_*, 8 for launch_dgemm_TN; 4 for the
 * packed families, 5 for their two-barrier modes).  The sgemm and
 * irr_thinB_sgemm launchers use the reference launchers' compiled-in
 * alpha = 5, beta = 3 (sgemm_release_V1.3.2/device_code/sgemm_launcher.c:11-12).
 * Asynchronous on `stream`.  BLASDSP_ERR_UNSUPPORTED for an unknown name. */
BLASDSP_API int blasdsp_group_exec(const char *name, int n_scalar, int n_ptr, const uint64_t *args,
                                   void *stream);

/* ---- per-mode BLAS entries with leading dimensions -------------------------
 * Generalisations of the packed kernels above: BLAS semantics
 * C = alpha op(A) op(B) + beta C and explicit leading dimensions, in dgemm_NN's
 * argument order.  blasdsp_sgemm_* run the AUTO dispatch (as blasdsp_sgemm);
 * blasdsp_irr_* the skinny kernel (the tile kernel where it does not fit). */
#define BLASDSP_DECLARE_MODE(name, T)                                                   \
    BLASDSP_API int name(int64_t lda, int64_t ldb, int64_t ldc, int64_t M, int64_t N, int64_t K,    \
             const T *A, const T *B, T *C, T alpha, T beta, void *stream);
BLASDSP_DECLARE_MODE(blasdsp_sgemm_NN, float)
BLASDSP_DECLARE_MODE(blasdsp_sgemm_NT, float)
BLASDSP_DECLARE_MODE(blasdsp_sgemm_TN, float)
BLASDSP_DECLARE_MODE(blasdsp_sgemm_TT, float)
BLASDSP_DECLARE_MODE(blasdsp_irr_shortA_dgemm_NN, double)
BLASDSP_DECLARE_MODE(blasdsp_irr_shortA_dgemm_NT, double)
BLASDSP_DECLARE_MODE(blasdsp_irr_shortA_dgemm_TN, double)
BLASDSP_DECLARE_MODE(blasdsp_irr_shortA_dgemm_TT, double)
BLASDSP_DECLARE_MODE(blasdsp_irr_shortA_sgemm_NN, float)
BLASDSP_DECLARE_MODE(blasdsp_irr_shortA_sgemm_NT, float)
BLASDSP_DECLARE_MODE(blasdsp_irr_shortA_sgemm_TN, float)
BLASDSP_DECLARE_MODE(blasdsp_irr_shortA_sgemm_TT, float)
BLASDSP_DECLARE_MODE(blasdsp_irr_thinB_dgemm_NN, double)
BLASDSP_DECLARE_MODE(blasdsp_irr_thinB_dgemm_NT, double)
BLASDSP_DECLARE_MODE(blasdsp_irr_thinB_dgemm_TN, double)
BLASDSP_DECLARE_MODE(blasdsp_irr_thinB_dgemm_TT, double)
BLASDSP_DECLARE_MODE(blasdsp_irr_thinB_sgemm_NN, float)
BLASDSP_DECLARE_MODE(blasdsp_irr_thinB_sgemm_NT, float)
BLASDSP_DECLARE_MODE(blasdsp_irr_thinB_sgemm_TN, float)
BLASDSP_DECLARE_MODE(blasdsp_irr_thinB_sgemm_TT, float)
#undef BLASDSP_DECLARE_MODE

/* ---- kernel selection (explicit) ----------------------------------------- */
enum blasdsp_algo {
    BLASDSP_ALGO_AUTO = 0,   /* shape-based dispatch (as blasdsp_dgemm)        */
    BLASDSP_ALGO_TILE = 1,   /* MFMA tile kernel                               */
    BLASDSP_ALGO_SKINNY = 2  /* bandwidth-bound skinny kernel (+ split-K)      */
};
/* BLASDSP_ALGO_SKINNY takes shapes whose small side is <= 64 and whose
 * leading dimensions keep the kernel's per-lane load offsets below 2^31
 * bytes (about 20 million elements for the long operand; about 1 million
 * FP32 / 0.5 million FP64 for a small operand stored with its short side
 * contiguous); other shapes run on the tile kernel.
 * splitk: K slices, their partial products added in a fixed order by a
 * second kernel (deterministic).  <= 0: chosen by the library — the tile
 * kernel splits K only when its C tiles fill the GPU's workgroup slots
 * poorly, the skinny kernel when its long side leaves them poorly filled. */
BLASDSP_API int blasdsp_dgemm_ex(char transA, char transB, int64_t M, int64_t N, int64_t K,
                     double alpha, const double *A, int64_t lda,
                     const double *B, int64_t ldb, double beta,
                     double *C, int64_t ldc, int algo, int splitk, void *stream);
BLASDSP_API int blasdsp_sgemm_ex(char transA, char transB, int64_t M, int64_t N, int64_t K,
                     float alpha, const float *A, int64_t lda,
                     const float *B, int64_t ldb, float beta,
                     float *C, int64_t ldc, int algo, int splitk, void *stream);


/* ---- precision forms (no counterpart in the reference) -------------------
 * FP32 GEMM on the BF16 matrix cores: each FP32 operand is split into BF16
 * pieces (hi + mid + lo for nterms = 6, hi + lo for nterms = 3) and the
 * products hh+hm+mh+hl+lh+mm (6) or hh+hl+lh (3) are accumulated in FP32 with
 * two-level (per-256-K-chunk) accumulation.  The pieces are staged in a
 * library workspace (about nterms/2 * 2 bytes per element of op(A) and op(B)).
 * Same semantics and arguments as blasdsp_sgemm; nterms must be 3 or 6. */
BLASDSP_API int blasdsp_sgemm_split_bf16(char transA, char transB, int64_t M, int64_t N, int64_t K,
                                         float alpha, const float *A, int64_t lda,
                                         const float *B, int64_t ldb, float beta,
                                         float *C, int64_t ldc, int nterms, void *stream);
/* FP32 GEMM on the FP16 matrix cores: every row of op(A) and column of op(B)
 * is scaled by a power of two (largest magnitude into [2^14, 2^15)), split
 * into FP16 hi + 2^-11 lo (22 significant bits), and the three products
 * hi*hi + 2^-11 (hi*lo + lo*hi) are accumulated in FP32.  Elements more than
 * 2^29 below their row's (column's) largest magnitude lose precision to FP16
 * subnormals.  Same semantics and arguments as blasdsp_sgemm. */
BLASDSP_API int blasdsp_sgemm_split_fp16(char transA, char transB, int64_t M, int64_t N, int64_t K,
                                         float alpha, const float *A, int64_t lda,
                                         const float *B, int64_t ldb, float beta,
                                         float *C, int64_t ldc, void *stream);

/* FP32 GEMM on the FP8 matrix cores (no counterpart in the reference): every
 * row of op(A) and column of op(B) is cut into blocks of 128 consecutive k,
 * each block gets the power-of-two scale 2^e with e the smallest exponent for
 * which amax 2^-e <= 448, and its elements are stored as E4M3 (round to
 * nearest even) of x 2^-e; the matrix cores apply the scales and accumulate
 * in FP32.  Precision is E4M3's (3 fraction bits): a normwise error of a few
 * 1e-2 on random inputs, reported, not bounded.  Non-finite inputs give NaN
 * in their rows and columns.  Same semantics and arguments as blasdsp_sgemm. */
BLASDSP_API int blasdsp_sgemm_fp8(char transA, char transB, int64_t M, int64_t N, int64_t K,
                                  float alpha, const float *A, int64_t lda,
                                  const float *B, int64_t ldb, float beta,
                                  float *C, int64_t ldc, void *stream);

/* ---- grouped execution (no counterpart in the reference) -----------------
 * Many GEMMs of uneven shapes, transpose modes and precisions in one
 * persistent launch per precision.  Each problem has the semantics of
 * blasdsp_sgemm / blasdsp_dgemm (alpha, beta are cast to float for FP32
 * problems) and produces the same bits as the tile kernel on its own. */
#define BLASDSP_FLOAT32 0
#define BLASDSP_FLOAT64 1
typedef struct blasdsp_gemm_desc {
    int dtype;              /* BLASDSP_FLOAT32 or BLASDSP_FLOAT64 */
    char transA, transB;    /* 'N' or 'T' */
    int64_t M, N, K;
    double alpha, beta;
    const void *A;
    int64_t lda;
    const void *B;
    int64_t ldb;
    void *C;
    int64_t ldc;
} blasdsp_gemm_desc;

/* Runs the whole group asynchronously on `stream` (a cached plan per stream). */
BLASDSP_API int blasdsp_gemm_grouped(const blasdsp_gemm_desc *descs, int count, void *stream);

/* A reusable plan: device problem list, pinned staging and instantiated
 * hipGraphs of the grouped kernels (one per set of problem lists a group can
 * fill).  run() uploads the list and replays, on `stream`, the graph of the
 * kernels the group needs; a run on another stream than the previous one
 * waits for the previous replay first.  A plan is used by one host thread at
 * a time. */
typedef struct blasdsp_group_plan blasdsp_group_plan;
BLASDSP_API int blasdsp_group_plan_create(int max_problems, blasdsp_group_plan **plan);
BLASDSP_API int blasdsp_group_plan_run(blasdsp_group_plan *plan, const blasdsp_gemm_desc *descs,
                                       int count, void *stream);
BLASDSP_API int blasdsp_group_plan_destroy(blasdsp_group_plan *plan);

/* ---- allocation-free GEMM --------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc): blasdsp_dgemm_ex / blasdsp_sgemm_ex (by desc->dtype)
 * with the same kernel, tile and K-slice choice, and so the same bits, but
 * the split-K partial sums go to a device array the library holds statically
 * (BLASDSP_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends)
 * instead of a hipMalloc'd workspace.  The call allocates nothing.
 * A slice count the library picks (splitk <= 0) is lowered until its partials
 * (slices x M x N elements) fit; an explicit splitk whose partials exceed the
 * array returns BLASDSP_ERR_NOMEM.  All calls on one GPU share its array, so
 * calls that overlap on one GPU (two streams or two host threads) race. */
#define BLASDSP_NOALLOC_SCRATCH_BYTES (64LL << 20)
BLASDSP_API int blasdsp_gemm_noalloc(const blasdsp_gemm_desc *desc, int algo, int splitk,
                                     void *stream);

/* ---- continuous-batching request loop (no counterpart in the reference) ----
 * blasdsp_serve() plays a list of GEMM requests against the GPU as a server
 * would: requests arrive at their (scaled) arrival times, are admitted FIFO
 * when their operands fit in an HBM budget (operands stay resident until the
 * request completes), and every admitted request not yet running is launched
 * as soon as a launch slot frees up — with grouped != 0 as one grouped
 * persistent launch (a hipGraph replay) per batch, else one GEMM launch per
 * request.  Operands are generated on the device: the stored A, B and C of a
 * request (dense, ld = stored columns) are the U[-1,1) generator streams of
 * seeds 3*seed, 3*seed+1 and 3*seed+2 (as blasdsp_fill_* with dist 0).
 * The call returns when every request has completed or was rejected. */
typedef struct blasdsp_request {
    int dtype;              /* BLASDSP_FLOAT32 or BLASDSP_FLOAT64 */
    char transA, transB;
    int64_t M, N, K;        /* all >= 1 */
    double alpha, beta;
    uint64_t seed;
    double arrival_s;       /* offered arrival time after the loop starts */
} blasdsp_request;

typedef struct blasdsp_request_result {
    double start_s;         /* admitted: operands allocated */
    double done_s;          /* completion observed by the loop */
    double latency_s;       /* done_s - scaled arrival time */
    double checksum;        /* FP64 sum of the elements of the result C */
    int batch;              /* launch the request ran in */
    int status;             /* BLASDSP_SUCCESS, or why the request was rejected */
} blasdsp_request_result;

typedef struct blasdsp_serve_config {
    int64_t memory_budget;  /* HBM bytes for resident operands */
    int max_batch;          /* requests per grouped launch, 1..16384 */
    int max_inflight;       /* launch slots (streams) in flight at once, 1..16 */
    int grouped;            /* 1: grouped launches; 0: one launch per request */
    double time_scale;      /* multiplies arrival times; 0: all arrive at once */
} blasdsp_serve_config;

typedef struct blasdsp_serve_stats {
    int64_t completed, rejected;
    double wall_s, gemms_per_s, tflops;
    double p50_s, p99_s, mean_s, max_s;     /* latency over completed requests */
    int64_t launches;
    double mean_batch;
    int64_t peak_resident_bytes;
} blasdsp_serve_stats;

/* outputs: NULL, or per-request host buffers receiving the dense M x N C. */
BLASDSP_API int blasdsp_serve(const blasdsp_request *reqs, int n, const blasdsp_serve_config *cfg,
                              blasdsp_request_result *results, blasdsp_serve_stats *stats,
                              void **outputs);

/* ---- autotuner (no counterpart in the reference) --------------------------
 * blasdsp_autotune() times the candidate kernels for one shape on `stream`
 * (the tile kernel with 64 and 128 tiles; when min(M, N) <= 64 also the
 * skinny kernel with split-K 1 .. 64) and records the fastest for the shape
 * class (precision, modes, round(log2) of M, N and K).  The AUTO dispatch of
 * every GEMM entry point and the grouped launch then use the recorded choice.
 * It allocates and frees temporary operands (it synchronises the device).
 * Tables persist as text through blasdsp_tune_save / _load. */
typedef struct blasdsp_tune_choice {
    int algo;               /* BLASDSP_ALGO_TILE or BLASDSP_ALGO_SKINNY */
    int tile;               /* tile edge (64 or 128) for the tile kernel */
    int splitk;             /* K slices (1: no split) */
    double time_us;         /* measured time of one call */
} blasdsp_tune_choice;
BLASDSP_API int blasdsp_autotune(int dtype, char transA, char transB, int64_t M, int64_t N, int64_t K,
                                 void *stream, blasdsp_tune_choice *out);
/* BLASDSP_ERR_UNSUPPORTED when the shape class has no entry. */
BLASDSP_API int blasdsp_tune_query(int dtype, char transA, char transB, int64_t M, int64_t N, int64_t K,
                                   blasdsp_tune_choice *out);
BLASDSP_API int blasdsp_tune_save(const char *path);
BLASDSP_API int blasdsp_tune_load(const char *path);
BLASDSP_API int blasdsp_tune_clear(void);

/* ---- transpose: B = A^T, A row-major M x N, B N x M ----------------------
 * HBM-bound (each element read once and written once).  The output must not
 * overlap the input.  Statuses as for the GEMMs.
 *   transpose/device_code/kernel.c:1167  transpose_16x16_dbuf(M, N, b_id, A, B)      FP64, packed
 *   transpose/device_code/kernel.c:2002  transpose_float_32x32_dbuf(M, N, b_id, A, B) FP32, packed
 *   transpose/device_code/kernel.c:1280,2110  transpose_null / transpose_float_null (empty launch)
 *   dgemm_release_V1.1/include/dgemm_kernel.h:8  transpose_A2B(lda, ldb, M, N, b_id, A, B)
 *   sgemm_release_V1.3.2/include/sgemm_kernel.h:8  transpose_A2B_float32(M, N, b_id, A, B)
 * Packed means lda = N and ldb = M.  Barrier ids are dropped; a stream is added. */
BLASDSP_API int transpose_16x16_dbuf(int64_t M, int64_t N, const double *A, double *B, void *stream);
BLASDSP_API int transpose_float_32x32_dbuf(int64_t M, int64_t N, const float *A, float *B, void *stream);
BLASDSP_API int transpose_null(int64_t M, int64_t N, const double *A, double *B, void *stream);
BLASDSP_API int transpose_float_null(int64_t M, int64_t N, const float *A, float *B, void *stream);
BLASDSP_API int transpose_A2B(int64_t lda, int64_t ldb, int64_t M, int64_t N, const double *A, double *B,
                              void *stream);
BLASDSP_API int transpose_A2B_float32(int64_t M, int64_t N, const float *A, float *B, void *stream);
/* Any leading dimensions (lda >= N, ldb >= M); no counterpart in the reference. */
BLASDSP_API int blasdsp_transpose_d(int64_t M, int64_t N, const double *A, int64_t lda, double *B,
                                    int64_t ldb, void *stream);
BLASDSP_API int blasdsp_transpose_s(int64_t M, int64_t N, const float *A, int64_t lda, float *B,
                                    int64_t ldb, void *stream);

/* ---- 2-D five-point stencil (one Jacobi sweep), FP64 ----------------------
 * A is an (M+2) x (N+2) grid with a one-element halo.  Every interior point
 * becomes (top + left + right + down + middle) * 0.2, summed left to right in
 * double, as the reference's host checker does
 * (stencil_2d_V1.2/host_code/matrix_stencil.c:33-47).  Out of place only: the
 * output must not overlap A.  HBM-bound.
 *   stencil_2d_V1.2/device_code/kernel_stencil.c:1042  stencil_vector(M, N, A, B)
 *       B is haloed like A, packed (ld N + 2); its interior is written and its
 *       halo left as it is (the layout of kernel_stencil.c:1095-1096)
 *   stencil_2d_V1.2/device_code/kernel_stencil.c:14    stencil_scalar(M, N, A, B)
 *       B is M x N, packed (ld N), as kernel_stencil.c:85-86 and the host checker
 *   stencil_2d_V1.2/device_code/kernel_stencil.c:1114  stencil_null(M, N, A, B)  (empty launch)
 * A stream is added. */
BLASDSP_API int stencil_vector(int64_t M, int64_t N, const double *A, double *B, void *stream);
BLASDSP_API int stencil_scalar(int64_t M, int64_t N, const double *A, double *B, void *stream);
BLASDSP_API int stencil_null(int64_t M, int64_t N, const double *A, double *B, void *stream);
/* Any leading dimensions (lda >= N + 2; ldb >= N + 2 with halo_out, else >= N);
 * no counterpart in the reference. */
BLASDSP_API int blasdsp_stencil(int64_t M, int64_t N, const double *A, int64_t lda, double *B, int64_t ldb,
                                int halo_out, void *stream);
/* `iters` Jacobi sweeps between two haloed grids X and Y (leading dimension
 * ld >= N + 2) whose halos both hold the boundary: sweep i reads X and writes
 * Y when i is even, the other way round when odd, so the last sweep's result
 * is in Y when iters is odd and in X when it is even.  The reference harness
 * means this loop (stencil_2d_V1.2/host_code/matrix_stencil.c:106-111), but
 * its pointer swap (switch_pointer, :28-33) exchanges copies of the pointers
 * and the launch arguments are fixed before the loop, so there every sweep
 * reads the same A. */
BLASDSP_API int blasdsp_stencil_iterate(int64_t M, int64_t N, double *X, double *Y, int64_t ld, int iters,
                                        void *stream);

/* ---- per-stream resources (no counterpart in the reference) ---------------
 * The library keeps, per (device, stream), a grow-only scratch buffer
 * (split-K partials, split and FP8 operand planes) and the cached plan of
 * blasdsp_gemm_grouped.  Call this before destroying a stream that was passed
 * to the library, with no library work pending on it, to free them; using
 * the stream again afterwards simply allocates anew. */
BLASDSP_API int blasdsp_stream_release(void *stream);

/* ---- diagnostics ---------------------------------------------------------
 * Streams `bytes` (a multiple of 16, 16-byte aligned) from device memory p
 * with `blocks` workgroups and writes one word per workgroup to out (device,
 * `blocks` words): the achievable HBM read bandwidth, the roof the
 * bandwidth-bound skinny kernels are measured against. */
BLASDSP_API int blasdsp_probe_read(const void *p, int64_t bytes, void *out, int blocks, void *stream);

/* The sustained rate of one matrix-core instruction form from registers
 * alone: `blocks` workgroups of 4 waves, each wave loading its operands once
 * from src (device, 16 KiB, 16-byte aligned: 1024 words of 16 bytes) and then
 * issuing `iters` rounds of independent MFMAs; out (device) receives
 * 256 * blocks floats.  Per wave and round: form 0 BF16 16x16x32 and 2 FP16
 * 16x16x32, 32 MFMAs; 1 BF16 32x32x16 and 4 FP32 32x32x2, 8 MFMAs; 3 FP32
 * 16x16x4, 32 MFMAs; 5 FP64 16x16x4 and 6 scaled FP8 16x16x128 (unit
 * scales), 16 MFMAs; 7 scaled FP8 32x32x64 (unit scales), 8 MFMAs.  The
 * caller times the launch. */
BLASDSP_API int blasdsp_probe_mfma(int form, const void *src, void *out, int blocks, int iters, void *stream);

/* Global -> LDS operand-load throughput from an L2-resident region: `blocks`
 * workgroups of 4 waves each load 32 KiB per iteration from src (device, 2 MiB)
 * into LDS, the previous iteration's loads still in flight; mode 0
 * direct-to-LDS buffer loads, 1 loads into registers and ds_write, 2 loads
 * into registers only.  out (device) receives 256 * blocks words.  The
 * caller times the launch (bytes = blocks * iters * 32 KiB). */
BLASDSP_API int blasdsp_probe_lds(int mode, const void *src, void *out, int blocks, int iters, void *stream);

/* ---- synthetic inputs ----------------------------------------------------
 * Fill the logical rows x cols matrix stored at X (device memory, leading
 * dimension ld, transposed storage when trans != 0) from the counter-based
 * generator: element (i, j) takes draw number i * cols + j of the seeded
 * stream, so the same logical matrix results for any ld or storage order.
 * Padding columns are set to `pad`.  dist: 0 uniform[-1,1), 1 normal(0,1),
 * 2 log-uniform magnitude 2^U[-8,8) with random sign, 3 (i*cols+j) % 11,
 * 4 (i*cols+j) % 7, 5 ones.  Uniform and the integer patterns are bit-exact
 * with the host generator; normal and log-uniform go through device
 * log/cos/exp2 and may differ from the host generator in the last bit. */
BLASDSP_API int blasdsp_fill_f64(double *X, int64_t rows, int64_t cols, int64_t ld, int trans,
                     int dist, uint64_t seed, double pad, void *stream);
BLASDSP_API int blasdsp_fill_f32(float *X, int64_t rows, int64_t cols, int64_t ld, int trans,
                     int dist, uint64_t seed, float pad, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* BLASDSP_H */
