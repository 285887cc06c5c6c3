/*
 * oracle.h — CPU oracle for the GEMM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it, and only as the checker (or, for cpu_baseline, as the timed
 * CPU restatement of the reference's host GEMM).
 *
 * Contents
 *   1. Counter-based generators for the seeded input distributions
 *      (uniform[-1,1], normal(0,1), log-uniform magnitude, the reference
 *      demos' integer patterns).
 *   2. Line-by-line restatements of the reference's host-side CPU GEMM
 *      (the CHECK_DATA loops of the demo harnesses), one function per
 *      published semantic, each citing the reference span it follows.
 *   3. The BLAS-epilogue variant of (2), which is the reported CPU baseline.
 *   4. An extended-precision (x87 long double accumulation) oracle for all
 *      four transpose modes with arbitrary leading dimensions.
 *
 * Matrices are row-major: element (i, j) of a rows x cols matrix with leading
 * dimension ld lives at X[i * ld + j].
 */
#ifndef BLASDSP_ORACLE_H
#define BLASDSP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 1. generators ------------------------------------------------------ */

enum oracle_dist {
    ORACLE_UNIFORM = 0,    /* i.i.d. uniform[-1, 1)                         */
    ORACLE_NORMAL = 1,     /* i.i.d. normal(0, 1) (Box-Muller)              */
    ORACLE_LOGUNIF = 2,    /* |x| = 2^U[-8, 8), random sign                 */
    ORACLE_PATTERN11 = 3,  /* x(i, j) = (i * cols + j) % 11  (demo matrix A) */
    ORACLE_PATTERN7 = 4,   /* x(i, j) = (i * cols + j) % 7   (demo matrix B) */
    ORACLE_ONES = 5        /* x = 1                                          */
};

uint64_t oracle_mix64(uint64_t x);
double oracle_value(int dist, uint64_t seed, uint64_t idx);

/* Fill the logical rows x cols matrix X.  Element (i, j) takes value
 * oracle_value(dist, seed, i * cols + j), so the same logical matrix results
 * for any ld and for transposed storage (trans != 0 stores (i, j) at
 * X[j * ld + i]).  Padding elements (beyond the logical row) are set to pad. */
void oracle_fill_f64(double *X, int64_t rows, int64_t cols, int64_t ld, int trans,
                     int dist, uint64_t seed, double pad);
void oracle_fill_f32(float *X, int64_t rows, int64_t cols, int64_t ld, int trans,
                     int dist, uint64_t seed, float pad);

/* ---- 2. restated reference host loops (NN layout only, as in the demos) -- */

/* C(i,j) += sum over K_G-chunks of an fma chain started at 0, then the tail
 * chunk (always added, possibly 0.0).  FP64.  No alpha/beta. */
void oracle_dgemm_v11_host(int64_t M, int64_t N, int64_t K,
                           const double *A, int64_t lda, const double *B, int64_t ldb,
                           double *C, int64_t ldc, int64_t K_G);

/* FP32 form of the same loop (float tmp, double fma), no alpha/beta. */
void oracle_sgemm_plain_host(int64_t M, int64_t N, int64_t K,
                             const float *A, int64_t lda, const float *B, int64_t ldb,
                             float *C, int64_t ldc, int64_t K_G);

/* sgemm V1.3.2: tmp = fma(alpha*A, B, tmp) per K_G chunk, C += tmp, C += beta. */
void oracle_sgemm_v132_host(int64_t M, int64_t N, int64_t K, float alpha, float beta,
                            const float *A, int64_t lda, const float *B, int64_t ldb,
                            float *C, int64_t ldc, int64_t K_G);

/* irr_thinB_sgemm V1.2: three-level K_G / K_A tree, alpha folded into B,
 * C += tmp_o + beta. */
void oracle_thinB_sgemm_v12_host(int64_t M, int64_t N, int64_t K, float alpha, float beta,
                                 const float *A, int64_t lda, const float *B, int64_t ldb,
                                 float *C, int64_t ldc, int64_t K_G, int64_t K_A);

/* ---- 3. BLAS-epilogue chunked loop (the CPU baseline) ------------------- */

/* acc = the K_G-chunked fma tree of (2); C = alpha * acc + beta * C
 * (beta == 0: C is not read).  NN layout. */
void oracle_dgemm_blas_host(int64_t M, int64_t N, int64_t K, double alpha,
                            const double *A, int64_t lda, const double *B, int64_t ldb,
                            double beta, double *C, int64_t ldc, int64_t K_G);
void oracle_sgemm_blas_host(int64_t M, int64_t N, int64_t K, float alpha,
                            const float *A, int64_t lda, const float *B, int64_t ldb,
                            float beta, float *C, int64_t ldc, int64_t K_G);

/* ---- 4. extended-precision oracle, all modes ---------------------------- */

/* C = alpha * op(A) op(B) + beta * C, long double accumulation, result
 * rounded once to double.  transA/transB are 'N' or 'T'.  op(A) is M x K:
 * A is M x K (lda >= K) for 'N', K x M (lda >= M) for 'T'.  Likewise op(B)
 * is K x N.  beta == 0: C is not read. */
void oracle_gemm_xp_f64(char transA, char transB, int64_t M, int64_t N, int64_t K,
                        double alpha, const double *A, int64_t lda,
                        const double *B, int64_t ldb, double beta,
                        double *C, int64_t ldc);
/* FP32 inputs, FP64 output C64 (ldc64), C32 supplies the beta term. */
void oracle_gemm_xp_f32(char transA, char transB, int64_t M, int64_t N, int64_t K,
                        double alpha, const float *A, int64_t lda,
                        const float *B, int64_t ldb, double beta,
                        const float *C32, int64_t ldc32, double *C64, int64_t ldc64);

/* Normwise relative error ||X - R||_F / ||R||_F over an M x N window, with
 * the max abs difference returned through maxabs (may be NULL). */
double oracle_normwise_f64(int64_t M, int64_t N, const double *X, int64_t ldx,
                           const double *R, int64_t ldr, double *maxabs);
double oracle_normwise_f32(int64_t M, int64_t N, const float *X, int64_t ldx,
                           const double *R, int64_t ldr, double *maxabs);

/* Y = op(A) X in extended precision (op(A) M x K, A stored K x M when trans;
 * X K x S and Y M x S row-major with leading dimension S, 1 <= S <= 16):
 * the random projection that checks whole GEMM results.  Returns 0, or -1 for
 * S out of range.  No reference counterpart. */
int oracle_project_f64(char trans, int64_t M, int64_t K, int64_t S, const double *A, int64_t lda,
                       const double *X, double *Y);
int oracle_project_f32(char trans, int64_t M, int64_t K, int64_t S, const float *A, int64_t lda,
                       const double *X, double *Y);

/* Scaled FP8 (E4M3, one power-of-two scale per 128-element block of each row
 * of op(A) and column of op(B)); see oracle.c for the rule.  No reference
 * counterpart. */
unsigned oracle_fp8_e4m3(float y);            /* RNE encode, |y| <= 448 */
float oracle_fp8_decode(unsigned q);
int oracle_fp8_block_exp(float amax);         /* smallest e with amax 2^-e <= 448 */
void oracle_sgemm_fp8(char transA, char transB, int64_t M, int64_t N, int64_t K, double alpha,
                      const float *A, int64_t lda, const float *B, int64_t ldb, double beta,
                      const float *C32, int64_t ldc32, double *C64, int64_t ldc64);

/* 2-D five-point stencil (restates reference_stencil,
 * stencil_2d_V1.2/host_code/matrix_stencil.c:33-47).  A: (M+2) x (N+2) with a
 * one-element halo, leading dimension lda >= N+2.  Output (row, col) of the
 * M x N interior goes to B[(row + off) * ldb + col + off]: off = 0 is the
 * packed layout, off = 1 the haloed one. */
void oracle_stencil(int64_t M, int64_t N, const double *A, int64_t lda, double *B, int64_t ldb, int off);

void oracle_set_threads(int n);
int oracle_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
