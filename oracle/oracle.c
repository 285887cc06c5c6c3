/*
 * oracle.c — CPU oracle for the GEMM hot path.  TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg;
 * never linked into or called by the product library.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).  Contraction is
 * off so that every expression rounds exactly as written, as the reference's
 * host code does when built with -std=c99 (sgemm_release_V1.3.2/host_code/Makefile:5).
 *
 * Parity pins (tests/test_oracle.py):
 *   - integer-pattern known answers (the reference demos' %11 / %7 matrices)
 *     against an exact int64 matrix product;
 *   - hand-derived rounding cases that fix the K_G / K_A chunk boundaries,
 *     the tail chunk and where alpha is folded;
 *   - exact rational (fractions.Fraction) brute force on tiny inputs for the
 *     long-double oracle and the normwise error;
 *   - closed forms (all-ones, identity, rank-1, K = 0) and mode invariance.
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define IDX(row, col, ld) ((int64_t)(row) * (int64_t)(ld) + (int64_t)(col))

static int g_threads = 0; /* 0: OpenMP default */

void oracle_set_threads(int n) { g_threads = n > 0 ? n : 0; }
int oracle_get_threads(void) { return g_threads > 0 ? g_threads : omp_get_max_threads(); }

/* ------------------------------------------------------------------------- */
/* 1. generators                                                             */
/* ------------------------------------------------------------------------- */

/* splitmix64 finaliser (Steele, Lea, Flood 2014). */
uint64_t oracle_mix64(uint64_t x)
{
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ULL;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBULL;
    x ^= x >> 31;
    return x;
}

/* Output number c of a splitmix64 stream whose state starts at a key derived
 * from the seed: counter-based, so any element can be produced on its own. */
static uint64_t stream_bits(uint64_t seed, uint64_t c)
{
    uint64_t key = oracle_mix64(seed * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL);
    return oracle_mix64(key + (c + 1) * 0x9E3779B97F4A7C15ULL);
}

/* [0, 1) with 53 random bits: m * 2^-53. */
static double unit_open_hi(uint64_t bits) { return (double)(bits >> 11) * 0x1.0p-53; }

double oracle_value(int dist, uint64_t seed, uint64_t idx)
{
    switch (dist) {
    case ORACLE_UNIFORM:
        /* 2u - 1 is exact in double. */
        return 2.0 * unit_open_hi(stream_bits(seed, idx)) - 1.0;
    case ORACLE_NORMAL: {
        double u1 = 1.0 - unit_open_hi(stream_bits(seed, 2 * idx));     /* (0, 1] */
        double u2 = unit_open_hi(stream_bits(seed, 2 * idx + 1));
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
    }
    case ORACLE_LOGUNIF: {
        uint64_t b0 = stream_bits(seed, 2 * idx);
        double u = unit_open_hi(stream_bits(seed, 2 * idx + 1));
        double mag = exp2(16.0 * u - 8.0);
        return (b0 >> 63) ? -mag : mag;
    }
    case ORACLE_PATTERN11:
        return (double)(idx % 11);
    case ORACLE_PATTERN7:
        return (double)(idx % 7);
    case ORACLE_ONES:
        return 1.0;
    default:
        return 0.0;
    }
}

void oracle_fill_f64(double *X, int64_t rows, int64_t cols, int64_t ld, int trans,
                     int dist, uint64_t seed, double pad)
{
    int64_t srows = trans ? cols : rows; /* stored rows */
    int64_t scols = trans ? rows : cols; /* stored logical columns */
#pragma omp parallel for num_threads(oracle_get_threads()) schedule(static)
    for (int64_t r = 0; r < srows; r++) {
        for (int64_t c = 0; c < ld; c++) {
            double v = pad;
            if (c < scols) {
                int64_t i = trans ? c : r, j = trans ? r : c;
                v = oracle_value(dist, seed, (uint64_t)(i * cols + j));
            }
            X[IDX(r, c, ld)] = v;
        }
    }
}

void oracle_fill_f32(float *X, int64_t rows, int64_t cols, int64_t ld, int trans,
                     int dist, uint64_t seed, float pad)
{
    int64_t srows = trans ? cols : rows;
    int64_t scols = trans ? rows : cols;
#pragma omp parallel for num_threads(oracle_get_threads()) schedule(static)
    for (int64_t r = 0; r < srows; r++) {
        for (int64_t c = 0; c < ld; c++) {
            float v = pad;
            if (c < scols) {
                int64_t i = trans ? c : r, j = trans ? r : c;
                v = (float)oracle_value(dist, seed, (uint64_t)(i * cols + j));
            }
            X[IDX(r, c, ld)] = v;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* 2. restated reference host loops                                          */
/* ------------------------------------------------------------------------- */

/* Follows dgemm_release_V1.1/host_code/dgemm_demo.c:233-252 (same loop in
 * irr_shortA_dgemm_release_V0.0/host_code/irr_shortA_dgemm_demo.c:205-224 and
 * irr_thinB_dgemm_release_V0.0/host_code/irr_thinB_dgemm_demo.c).  Per
 * element: every full K_G chunk is an fma chain from 0.0 added into C, then
 * the tail chunk k in [K - K%K_G, K) likewise, added even when empty. */
void oracle_dgemm_v11_host(int64_t M, int64_t N, int64_t K,
                           const double *A, int64_t lda, const double *B, int64_t ldb,
                           double *C, int64_t ldc, int64_t K_G)
{
#pragma omp parallel for num_threads(oracle_get_threads()) collapse(2) schedule(static)
    for (int64_t i = 0; i < M; i++) {
        for (int64_t j = 0; j < N; j++) {
            for (int64_t k = 0; k < K - K % K_G; k += K_G) {
                double tmp = 0.0;
                for (int64_t kk = 0; kk < K_G; kk++)
                    tmp = fma(A[IDX(i, k + kk, lda)], B[IDX(k + kk, j, ldb)], tmp);
                C[IDX(i, j, ldc)] += tmp;
            }
            double tmp = 0.0;
            for (int64_t k = K - K % K_G; k < K; k++)
                tmp = fma(A[IDX(i, k, lda)], B[IDX(k, j, ldb)], tmp);
            C[IDX(i, j, ldc)] += tmp;
        }
    }
}

/* Follows irr_shortA_sgemm_release_V0.0/host_code/irr_shortA_sgemm_demo.c:205-224:
 * the loop above with dataType float — tmp is a float, fma() runs in double
 * and its result is rounded to float on assignment. */
void oracle_sgemm_plain_host(int64_t M, int64_t N, int64_t K,
                             const float *A, int64_t lda, const float *B, int64_t ldb,
                             float *C, int64_t ldc, int64_t K_G)
{
#pragma omp parallel for num_threads(oracle_get_threads()) collapse(2) schedule(static)
    for (int64_t i = 0; i < M; i++) {
        for (int64_t j = 0; j < N; j++) {
            for (int64_t k = 0; k < K - K % K_G; k += K_G) {
                float tmp = 0.0f;
                for (int64_t kk = 0; kk < K_G; kk++)
                    tmp = (float)fma((double)A[IDX(i, k + kk, lda)],
                                     (double)B[IDX(k + kk, j, ldb)], (double)tmp);
                C[IDX(i, j, ldc)] += tmp;
            }
            float tmp = 0.0f;
            for (int64_t k = K - K % K_G; k < K; k++)
                tmp = (float)fma((double)A[IDX(i, k, lda)], (double)B[IDX(k, j, ldb)],
                                 (double)tmp);
            C[IDX(i, j, ldc)] += tmp;
        }
    }
}

/* Follows sgemm_release_V1.3.2/host_code/sgemm_demo.c:214-242.  alpha * A is
 * a float product (rounded to float) before it enters the double fma; after
 * the chunks, the scalar beta is added (the reference's C_plus_beta semantic,
 * sgemm_release_V1.3.2/device_code/sgemm_kernel.c:25-160). */
void oracle_sgemm_v132_host(int64_t M, int64_t N, int64_t K, float alpha, float beta,
                            const float *A, int64_t lda, const float *B, int64_t ldb,
                            float *C, int64_t ldc, int64_t K_G)
{
#pragma omp parallel for num_threads(oracle_get_threads()) collapse(2) schedule(static)
    for (int64_t i = 0; i < M; i++) {
        for (int64_t j = 0; j < N; j++) {
            for (int64_t k = 0; k < K - K % K_G; k += K_G) {
                float tmp = 0.0f;
                for (int64_t kk = 0; kk < K_G; kk++) {
                    float a = alpha * A[IDX(i, k + kk, lda)];
                    tmp = (float)fma((double)a, (double)B[IDX(k + kk, j, ldb)], (double)tmp);
                }
                C[IDX(i, j, ldc)] += tmp;
            }
            float tmp = 0.0f;
            for (int64_t k = K - K % K_G; k < K; k++) {
                float a = alpha * A[IDX(i, k, lda)];
                tmp = (float)fma((double)a, (double)B[IDX(k, j, ldb)], (double)tmp);
            }
            C[IDX(i, j, ldc)] += tmp;
            C[IDX(i, j, ldc)] += beta;
        }
    }
}

static int64_t min64(int64_t a, int64_t b) { return a < b ? a : b; }

/* Follows irr_thinB_sgemm_release_V1.2/host_code/irr_thinB_sgemm_demo.c:217-240:
 * a three-level tree (K_G outer chunks of K_A inner chunks, each inner chunk an
 * fma chain from 0), alpha folded into the B operand, and C += tmp_o + beta. */
void oracle_thinB_sgemm_v12_host(int64_t M, int64_t N, int64_t K, float alpha, float beta,
                                 const float *A, int64_t lda, const float *B, int64_t ldb,
                                 float *C, int64_t ldc, int64_t K_G, int64_t K_A)
{
#pragma omp parallel for num_threads(oracle_get_threads()) collapse(2) schedule(static)
    for (int64_t i = 0; i < M; i++) {
        for (int64_t j = 0; j < N; j++) {
            float tmp_o = 0.0f;
            for (int64_t k_o = 0; k_o < K; k_o += K_G) {
                int64_t kg = min64(K_G, K - k_o);
                float tmp_i = 0.0f;
                for (int64_t k_i = 0; k_i < kg; k_i += K_A) {
                    int64_t ka = min64(K_A, kg - k_i);
                    float tmp_s = 0.0f;
                    for (int64_t k_s = 0; k_s < ka; k_s++) {
                        int64_t k = k_o + k_i + k_s;
                        float b = alpha * B[IDX(k, j, ldb)];
                        tmp_s = (float)fma((double)A[IDX(i, k, lda)], (double)b, (double)tmp_s);
                    }
                    tmp_i += tmp_s;
                }
                tmp_o += tmp_i;
            }
            C[IDX(i, j, ldc)] += tmp_o + beta;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* 3. BLAS-epilogue chunked loop = CPU baseline                              */
/* ------------------------------------------------------------------------- */

/* The accumulation tree of dgemm_demo.c:233-252 (K_G chunks, fma chain from
 * 0, chunks summed in order), finished with the BLAS epilogue
 * C = alpha * acc + beta * C that the product entry points implement. */
void oracle_dgemm_blas_host(int64_t M, int64_t N, int64_t K, double alpha,
                            const double *A, int64_t lda, const double *B, int64_t ldb,
                            double beta, double *C, int64_t ldc, int64_t K_G)
{
#pragma omp parallel for num_threads(oracle_get_threads()) collapse(2) schedule(static)
    for (int64_t i = 0; i < M; i++) {
        for (int64_t j = 0; j < N; j++) {
            double acc = 0.0;
            for (int64_t k = 0; k < K - K % K_G; k += K_G) {
                double tmp = 0.0;
                for (int64_t kk = 0; kk < K_G; kk++)
                    tmp = fma(A[IDX(i, k + kk, lda)], B[IDX(k + kk, j, ldb)], tmp);
                acc += tmp;
            }
            double tmp = 0.0;
            for (int64_t k = K - K % K_G; k < K; k++)
                tmp = fma(A[IDX(i, k, lda)], B[IDX(k, j, ldb)], tmp);
            acc += tmp;
            double out = alpha * acc;
            if (beta != 0.0) out = out + beta * C[IDX(i, j, ldc)];
            C[IDX(i, j, ldc)] = out;
        }
    }
}

/* FP32 form (float tmp, double fma, as sgemm_demo.c:214-242 without alpha
 * folding), BLAS epilogue in float. */
void oracle_sgemm_blas_host(int64_t M, int64_t N, int64_t K, float alpha,
                            const float *A, int64_t lda, const float *B, int64_t ldb,
                            float beta, float *C, int64_t ldc, int64_t K_G)
{
#pragma omp parallel for num_threads(oracle_get_threads()) collapse(2) schedule(static)
    for (int64_t i = 0; i < M; i++) {
        for (int64_t j = 0; j < N; j++) {
            float acc = 0.0f;
            for (int64_t k = 0; k < K - K % K_G; k += K_G) {
                float tmp = 0.0f;
                for (int64_t kk = 0; kk < K_G; kk++)
                    tmp = (float)fma((double)A[IDX(i, k + kk, lda)],
                                     (double)B[IDX(k + kk, j, ldb)], (double)tmp);
                acc += tmp;
            }
            float tmp = 0.0f;
            for (int64_t k = K - K % K_G; k < K; k++)
                tmp = (float)fma((double)A[IDX(i, k, lda)], (double)B[IDX(k, j, ldb)],
                                 (double)tmp);
            acc += tmp;
            float out = alpha * acc;
            if (beta != 0.0f) out = out + beta * C[IDX(i, j, ldc)];
            C[IDX(i, j, ldc)] = out;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* 4. extended-precision oracle                                              */
/* ------------------------------------------------------------------------- */

/* op(B) packed column-major (column j contiguous over k) so the inner loop
 * streams; accumulation in x87 long double (64-bit significand). */
void oracle_gemm_xp_f64(char transA, char transB, int64_t M, int64_t N, int64_t K,
                        double alpha, const double *A, int64_t lda,
                        const double *B, int64_t ldb, double beta,
                        double *C, int64_t ldc)
{
    int ta = (transA == 'T' || transA == 't');
    int tb = (transB == 'T' || transB == 't');
    double *Bp = (double *)malloc(sizeof(double) * (size_t)(K > 0 ? K : 1) * (size_t)(N > 0 ? N : 1));
    if (!Bp) return;
#pragma omp parallel for num_threads(oracle_get_threads()) schedule(static)
    for (int64_t j = 0; j < N; j++)
        for (int64_t k = 0; k < K; k++)
            Bp[IDX(j, k, K)] = tb ? B[IDX(j, k, ldb)] : B[IDX(k, j, ldb)];
#pragma omp parallel num_threads(oracle_get_threads())
    {
        double *arow = (double *)malloc(sizeof(double) * (size_t)(K > 0 ? K : 1));
#pragma omp for schedule(static)
        for (int64_t i = 0; i < M; i++) {
            for (int64_t k = 0; k < K; k++)
                arow[k] = ta ? A[IDX(k, i, lda)] : A[IDX(i, k, lda)];
            for (int64_t j = 0; j < N; j++) {
                const double *bc = Bp + IDX(j, 0, K);
                long double acc = 0.0L;
                for (int64_t k = 0; k < K; k++) acc += (long double)arow[k] * (long double)bc[k];
                long double out = (long double)alpha * acc;
                if (beta != 0.0) out += (long double)beta * (long double)C[IDX(i, j, ldc)];
                C[IDX(i, j, ldc)] = (double)out;
            }
        }
        free(arow);
    }
    free(Bp);
}

void oracle_gemm_xp_f32(char transA, char transB, int64_t M, int64_t N, int64_t K,
                        double alpha, const float *A, int64_t lda,
                        const float *B, int64_t ldb, double beta,
                        const float *C32, int64_t ldc32, double *C64, int64_t ldc64)
{
    int ta = (transA == 'T' || transA == 't');
    int tb = (transB == 'T' || transB == 't');
    double *Bp = (double *)malloc(sizeof(double) * (size_t)(K > 0 ? K : 1) * (size_t)(N > 0 ? N : 1));
    if (!Bp) return;
#pragma omp parallel for num_threads(oracle_get_threads()) schedule(static)
    for (int64_t j = 0; j < N; j++)
        for (int64_t k = 0; k < K; k++)
            Bp[IDX(j, k, K)] = tb ? (double)B[IDX(j, k, ldb)] : (double)B[IDX(k, j, ldb)];
#pragma omp parallel num_threads(oracle_get_threads())
    {
        double *arow = (double *)malloc(sizeof(double) * (size_t)(K > 0 ? K : 1));
#pragma omp for schedule(static)
        for (int64_t i = 0; i < M; i++) {
            for (int64_t k = 0; k < K; k++)
                arow[k] = ta ? (double)A[IDX(k, i, lda)] : (double)A[IDX(i, k, lda)];
            for (int64_t j = 0; j < N; j++) {
                const double *bc = Bp + IDX(j, 0, K);
                long double acc = 0.0L;
                for (int64_t k = 0; k < K; k++) acc += (long double)arow[k] * (long double)bc[k];
                long double out = (long double)alpha * acc;
                if (beta != 0.0) out += (long double)beta * (long double)C32[IDX(i, j, ldc32)];
                C64[IDX(i, j, ldc64)] = (double)out;
            }
        }
        free(arow);
    }
    free(Bp);
}

double oracle_normwise_f64(int64_t M, int64_t N, const double *X, int64_t ldx,
                           const double *R, int64_t ldr, double *maxabs)
{
    long double num = 0.0L, den = 0.0L;
    double mx = 0.0;
    for (int64_t i = 0; i < M; i++)
        for (int64_t j = 0; j < N; j++) {
            long double r = R[IDX(i, j, ldr)];
            long double d = (long double)X[IDX(i, j, ldx)] - r;
            num += d * d;
            den += r * r;
            double ad = fabs((double)d);
            if (ad > mx || ad != ad) mx = ad;
        }
    if (maxabs) *maxabs = mx;
    if (den == 0.0L) return num == 0.0L ? 0.0 : INFINITY;
    return (double)sqrtl(num / den);
}

double oracle_normwise_f32(int64_t M, int64_t N, const float *X, int64_t ldx,
                           const double *R, int64_t ldr, double *maxabs)
{
    long double num = 0.0L, den = 0.0L;
    double mx = 0.0;
    for (int64_t i = 0; i < M; i++)
        for (int64_t j = 0; j < N; j++) {
            long double r = R[IDX(i, j, ldr)];
            long double d = (long double)X[IDX(i, j, ldx)] - r;
            num += d * d;
            den += r * r;
            double ad = fabs((double)d);
            if (ad > mx || ad != ad) mx = ad;
        }
    if (maxabs) *maxabs = mx;
    if (den == 0.0L) return num == 0.0L ? 0.0 : INFINITY;
    return (double)sqrtl(num / den);
}

/* ---- random projection -----------------------------------------------------
 * Y = op(A) X with long-double sums, rounded once: op(A) is M x K (A stored
 * K x M when trans), X is K x S and Y is M x S, both row-major with leading
 * dimension S.  The checker for whole GEMM results: a result G of
 * R = alpha op(A) op(B) + beta C0 is compared through G X against
 * R X = alpha op(A) (op(B) X) + beta C0 X for random X (Freivalds' method),
 * O((M + K) N S) work where the full product would cost O(M N K). */
#define PROJ_MAX_S 16
#define PROJ_COLS 32

static inline long double proj_elem(const void *A, int is32, int64_t idx)
{
    return is32 ? (long double)((const float *)A)[idx] : (long double)((const double *)A)[idx];
}

static int project(int trans, int64_t M, int64_t K, int64_t S, const void *A, int is32, int64_t lda,
                   const double *X, double *Y)
{
    if (S < 1 || S > PROJ_MAX_S || M < 0 || K < 0) return -1;
    if (!trans) {
#pragma omp parallel for num_threads(oracle_get_threads()) schedule(static)
        for (int64_t i = 0; i < M; i++) {
            long double acc[PROJ_MAX_S] = {0.0L};
            for (int64_t k = 0; k < K; k++) {
                const long double a = proj_elem(A, is32, IDX(i, k, lda));
                for (int64_t s = 0; s < S; s++) acc[s] += a * (long double)X[IDX(k, s, S)];
            }
            for (int64_t s = 0; s < S; s++) Y[IDX(i, s, S)] = (double)acc[s];
        }
        return 0;
    }
    /* stored K x M: each thread takes PROJ_COLS consecutive columns and walks
     * down the rows, so every row segment it reads is contiguous */
    const int64_t nblk = (M + PROJ_COLS - 1) / PROJ_COLS;
#pragma omp parallel for num_threads(oracle_get_threads()) schedule(dynamic)
    for (int64_t b = 0; b < nblk; b++) {
        const int64_t i0 = b * PROJ_COLS;
        const int64_t nc = M - i0 < PROJ_COLS ? M - i0 : PROJ_COLS;
        long double acc[PROJ_COLS][PROJ_MAX_S];
        for (int64_t c = 0; c < nc; c++)
            for (int64_t s = 0; s < S; s++) acc[c][s] = 0.0L;
        for (int64_t k = 0; k < K; k++)
            for (int64_t c = 0; c < nc; c++) {
                const long double a = proj_elem(A, is32, IDX(k, i0 + c, lda));
                for (int64_t s = 0; s < S; s++) acc[c][s] += a * (long double)X[IDX(k, s, S)];
            }
        for (int64_t c = 0; c < nc; c++)
            for (int64_t s = 0; s < S; s++) Y[IDX(i0 + c, s, S)] = (double)acc[c][s];
    }
    return 0;
}

int oracle_project_f64(char trans, int64_t M, int64_t K, int64_t S, const double *A, int64_t lda,
                       const double *X, double *Y)
{
    return project(trans == 'T' || trans == 't', M, K, S, A, 0, lda, X, Y);
}

int oracle_project_f32(char trans, int64_t M, int64_t K, int64_t S, const float *A, int64_t lda,
                       const double *X, double *Y)
{
    return project(trans == 'T' || trans == 't', M, K, S, A, 1, lda, X, Y);
}

/* ---- scaled FP8 (E4M3) GEMM ------------------------------------------------
 * No reference counterpart: the FP8 form of the FP32 GEMM.  The rule, shared
 * with the library's quantizer:
 *   - every row of op(A) and every column of op(B) is cut into blocks of 128
 *     consecutive k (the last block zero-padded);
 *   - a block with largest magnitude amax = f 2^x (f in [0.5, 1)) gets the
 *     exponent e = x - 9 + (f > 0.875), the smallest e with amax 2^-e <= 448
 *     (the largest finite E4M3 value), clamped to [-127, 127]; a zero block
 *     gets e = 0;
 *   - each element is stored as q = E4M3(x 2^-e), rounded to nearest even
 *     (OCP E4M3: bias 7, subnormal step 2^-9, no infinities);
 *   - C = alpha sum_blocks 2^(ea + eb) sum_k qa qb + beta C.
 * Pinned in tests/test_oracle.py: the encoder against PyTorch's float8_e4m3fn
 * conversion, the block exponent against its defining inequality, and the
 * quantized product against exact rational arithmetic on small inputs. */

unsigned oracle_fp8_e4m3(float y)
{
    unsigned sign = signbit(y) ? 0x80u : 0u;
    float a = fabsf(y);
    if (a != a) return 0x7Fu;
    if (a < 0x1p-6f) {                      /* subnormal range: multiples of 2^-9 */
        float m = rintf(a * 0x1p9f);        /* 0 .. 8; 8 is the smallest normal */
        return sign | (unsigned)m;
    }
    int x;
    (void)frexpf(a, &x);                    /* a = f 2^x, f in [0.5, 1) */
    int E = x - 1;                          /* a in [2^E, 2^(E+1)) */
    float m = rintf(ldexpf(a, 3 - E));      /* 8 .. 16 */
    if (m >= 16.0f) {
        m = 8.0f;
        E += 1;
    }
    if (E > 8 || (E == 8 && m > 14.0f)) return sign | 0x7Eu; /* saturate to 448 */
    return sign | (unsigned)((E + 7) << 3) | ((unsigned)m - 8u);
}

float oracle_fp8_decode(unsigned q)
{
    const float s = (q & 0x80u) ? -1.0f : 1.0f;
    const unsigned e = (q >> 3) & 15u, m = q & 7u;
    if (e == 15u && m == 7u) return NAN;
    if (e == 0u) return s * ldexpf((float)m, -9);
    return s * ldexpf((float)(8u + m), (int)e - 10);
}

int oracle_fp8_block_exp(float amax)
{
    if (!(amax > 0.0f)) return 0;
    int x;
    const float f = frexpf(amax, &x);
    int e = x - 9 + (f > 0.875f ? 1 : 0);
    if (e < -127) e = -127;
    if (e > 127) e = 127;
    return e;
}

/* op(X) row r (r < R), k-th element, X stored row-major with ld; trans: op(X) = X^T. */
static float opx(const float *X, int64_t ld, int trans, int64_t r, int64_t k)
{
    return trans ? X[IDX(k, r, ld)] : X[IDX(r, k, ld)];
}

/* Quantize the R vectors of length K of op(X) (rows of op(A): trans = ta;
 * columns of op(B): pass B with the opposite sense) into q[R][KB*128] codes
 * decoded to float, and their block exponents e[R][KB]. */
static void fp8_quantize(const float *X, int64_t ld, int trans, int64_t R, int64_t K, float *q, int *e)
{
    const int64_t KB = (K + 127) / 128;
#pragma omp parallel for num_threads(oracle_get_threads()) schedule(static)
    for (int64_t r = 0; r < R; r++)
        for (int64_t b = 0; b < KB; b++) {
            float amax = 0.0f;
            for (int64_t k = b * 128; k < K && k < b * 128 + 128; k++) amax = fmaxf(amax, fabsf(opx(X, ld, trans, r, k)));
            const int be = oracle_fp8_block_exp(amax);
            e[IDX(r, b, KB)] = be;
            for (int64_t k = b * 128; k < b * 128 + 128; k++) {
                const float v = k < K ? ldexpf(opx(X, ld, trans, r, k), -be) : 0.0f;
                q[IDX(r, k, KB * 128)] = oracle_fp8_decode(oracle_fp8_e4m3(v));
            }
        }
}

void oracle_sgemm_fp8(char transA, char transB, int64_t M, int64_t N, int64_t K, double alpha,
                      const float *A, int64_t lda, const float *B, int64_t ldb, double beta,
                      const float *C32, int64_t ldc32, double *C64, int64_t ldc64)
{
    const int ta = (transA == 'T' || transA == 't');
    const int tb = (transB == 'T' || transB == 't');
    const int64_t KB = (K + 127) / 128, Kp = KB * 128;
    float *qa = (float *)malloc(sizeof(float) * (size_t)(M > 0 ? M : 1) * (size_t)(Kp > 0 ? Kp : 1));
    float *qb = (float *)malloc(sizeof(float) * (size_t)(N > 0 ? N : 1) * (size_t)(Kp > 0 ? Kp : 1));
    int *ea = (int *)malloc(sizeof(int) * (size_t)(M > 0 ? M : 1) * (size_t)(KB > 0 ? KB : 1));
    int *eb = (int *)malloc(sizeof(int) * (size_t)(N > 0 ? N : 1) * (size_t)(KB > 0 ? KB : 1));
    if (!qa || !qb || !ea || !eb) {
        free(qa); free(qb); free(ea); free(eb);
        return;
    }
    fp8_quantize(A, lda, ta, M, K, qa, ea);
    /* column j of op(B): op(B)[k][j] = tb ? B[j][k] : B[k][j], i.e. row j of the
     * transposed view, which is "trans" exactly when tb is false */
    fp8_quantize(B, ldb, !tb, N, K, qb, eb);
#pragma omp parallel for num_threads(oracle_get_threads()) schedule(static)
    for (int64_t i = 0; i < M; i++)
        for (int64_t j = 0; j < N; j++) {
            long double acc = 0.0L;
            for (int64_t b = 0; b < KB; b++) {
                long double part = 0.0L;
                for (int64_t k = b * 128; k < b * 128 + 128; k++)
                    part += (long double)qa[IDX(i, k, Kp)] * (long double)qb[IDX(j, k, Kp)];
                acc += ldexpl(part, ea[IDX(i, b, KB)] + eb[IDX(j, b, KB)]);
            }
            long double out = (long double)alpha * acc;
            if (beta != 0.0) out += (long double)beta * (long double)C32[IDX(i, j, ldc32)];
            C64[IDX(i, j, ldc64)] = (double)out;
        }
    free(qa); free(qb); free(ea); free(eb);
}

/* ------------------------------------------------------------------------- */
/* 2-D five-point stencil                                                    */
/* ------------------------------------------------------------------------- */

/* reference_stencil (stencil_2d_V1.2/host_code/matrix_stencil.c:33-47):
 *   top    = A[row][col + 1]      left  = A[row + 1][col]
 *   right  = A[row + 1][col + 2]  down  = A[row + 2][col + 1]
 *   middle = A[row + 1][col + 1]
 *   out    = (top + left + right + down + middle) * 0.2
 * in haloed coordinates, the sum taken left to right in double and then
 * multiplied by the double nearest 0.2.  (The reference's OpenMP loop shares
 * its five temporaries between threads, a data race; the restatement keeps
 * them per element, which is what the loop means.)  The reference writes the
 * packed M x N layout (off = 0, as stencil_scalar does,
 * stencil_2d_V1.2/device_code/kernel_stencil.c:85-86); stencil_vector writes
 * the interior of a haloed (M+2) x (N+2) B (off = 1, kernel_stencil.c:1095-1096). */
void oracle_stencil(int64_t M, int64_t N, const double *A, int64_t lda, double *B, int64_t ldb, int off)
{
#pragma omp parallel for num_threads(oracle_get_threads()) schedule(static)
    for (int64_t row = 0; row < M; row++)
        for (int64_t col = 0; col < N; col++) {
            const double top = A[IDX(row, col + 1, lda)];
            const double left = A[IDX(row + 1, col, lda)];
            const double right = A[IDX(row + 1, col + 2, lda)];
            const double down = A[IDX(row + 2, col + 1, lda)];
            const double middle = A[IDX(row + 1, col + 1, lda)];
            B[IDX(row + off, col + off, ldb)] = (top + left + right + down + middle) * 0.2;
        }
}
