/* gemm_demo.c — the reference's demo harness, on this library.
 *
 * Same command line and flow as the reference's host demos
 * (dgemm_release_V1.1/host_code/dgemm_demo.c:26-200 and its siblings for
 * sgemm, irr_shortA_* and irr_thinB_*):
 *
 *   gemm_demo <family> dev_id M K N [num_threads]
 *
 * family: dgemm, sgemm, irr_shortA_dgemm, irr_shortA_sgemm, irr_thinB_dgemm,
 * irr_thinB_sgemm.  Operands are allocated with the reference's padding
 * (rows + 64, K + 128, columns + 256) and leading dimensions, A and B also as
 * stored transposes, filled on the device (U[-1,1), the same logical matrix in
 * both storage orders).  The empty kernel is timed first and subtracted, as
 * the reference does (dgemm_demo.c:140-158), then each of NN, NT, TN, TT runs
 * NUM_TESTS times from C = 0 and prints its time and GFLOPS.  num_threads (DSP
 * cores there) is accepted and ignored.
 *
 * Check: the four modes read the same logical matrices, so their results
 * must agree: the tile kernel accumulates every element in the same order in
 * every mode (bit-identical), the skinny kernel groups k differently for
 * K-contiguous and k-major streamed operands (agreement to rounding).  The
 * program prints the largest normwise difference from NN and exits 1 above
 * 1e-12 (FP64) / 1e-5 (FP32).
 */
#include <math.h>
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "blasdsp.h"

#define NUM_TESTS 5

#define HIP_OK(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            fprintf(stderr, "HIP error %d at %s:%d\n", (int)e_, __FILE__, __LINE__);   \
            exit(3);                                                                   \
        }                                                                              \
    } while (0)

#define BLAS_OK(x)                                                                     \
    do {                                                                               \
        int s_ = (x);                                                                  \
        if (s_ != BLASDSP_SUCCESS) {                                                   \
            fprintf(stderr, "%s at %s:%d\n", blasdsp_status_string(s_), __FILE__, __LINE__); \
            exit(2);                                                                   \
        }                                                                              \
    } while (0)

static double timer_us(void)
{
    struct timeval t;
    gettimeofday(&t, NULL);
    return (double)t.tv_sec * 1e6 + (double)t.tv_usec;
}

typedef int (*gemm_fn)(int64_t, int64_t, int64_t, int64_t, int64_t, int64_t, const void *, const void *, void *,
                       double, double, void *);

/* Adapters with one signature over the library's typed entry points. */
#define ADAPT_D(name)                                                                                  \
    static int a_##name(int64_t lda, int64_t ldb, int64_t ldc, int64_t M, int64_t N, int64_t K,        \
                        const void *A, const void *B, void *C, double al, double be, void *s)          \
    {                                                                                                  \
        return name(lda, ldb, ldc, M, N, K, (const double *)A, (const double *)B, (double *)C, al, be, s); \
    }
#define ADAPT_S(name)                                                                                  \
    static int a_##name(int64_t lda, int64_t ldb, int64_t ldc, int64_t M, int64_t N, int64_t K,        \
                        const void *A, const void *B, void *C, double al, double be, void *s)          \
    {                                                                                                  \
        return name(lda, ldb, ldc, M, N, K, (const float *)A, (const float *)B, (float *)C, (float)al,  \
                    (float)be, s);                                                                      \
    }
ADAPT_D(dgemm_NN) ADAPT_D(dgemm_NT) ADAPT_D(dgemm_TN) ADAPT_D(dgemm_TT) ADAPT_D(dgemm_null)
ADAPT_S(blasdsp_sgemm_NN) ADAPT_S(blasdsp_sgemm_NT) ADAPT_S(blasdsp_sgemm_TN) ADAPT_S(blasdsp_sgemm_TT)
ADAPT_D(blasdsp_irr_shortA_dgemm_NN) ADAPT_D(blasdsp_irr_shortA_dgemm_NT) ADAPT_D(blasdsp_irr_shortA_dgemm_TN)
ADAPT_D(blasdsp_irr_shortA_dgemm_TT)
ADAPT_S(blasdsp_irr_shortA_sgemm_NN) ADAPT_S(blasdsp_irr_shortA_sgemm_NT) ADAPT_S(blasdsp_irr_shortA_sgemm_TN)
ADAPT_S(blasdsp_irr_shortA_sgemm_TT)
ADAPT_D(blasdsp_irr_thinB_dgemm_NN) ADAPT_D(blasdsp_irr_thinB_dgemm_NT) ADAPT_D(blasdsp_irr_thinB_dgemm_TN)
ADAPT_D(blasdsp_irr_thinB_dgemm_TT)
ADAPT_S(blasdsp_irr_thinB_sgemm_NN) ADAPT_S(blasdsp_irr_thinB_sgemm_NT) ADAPT_S(blasdsp_irr_thinB_sgemm_TN)
ADAPT_S(blasdsp_irr_thinB_sgemm_TT)

/* sgemm_null takes the reference's packed list (M, N, K, A, B, C, alpha, beta) */
static int a_sgemm_null(int64_t lda, int64_t ldb, int64_t ldc, int64_t M, int64_t N, int64_t K, const void *A,
                        const void *B, void *C, double al, double be, void *s)
{
    (void)lda; (void)ldb; (void)ldc;
    return sgemm_null(M, N, K, (const float *)A, (const float *)B, (float *)C, (float)al, (float)be, s);
}

struct family {
    const char *name;
    int f64;
    gemm_fn mode[4]; /* NN, NT, TN, TT */
    gemm_fn null_fn;
};

static const struct family FAMILIES[] = {
    {"dgemm", 1, {a_dgemm_NN, a_dgemm_NT, a_dgemm_TN, a_dgemm_TT}, a_dgemm_null},
    {"sgemm", 0, {a_blasdsp_sgemm_NN, a_blasdsp_sgemm_NT, a_blasdsp_sgemm_TN, a_blasdsp_sgemm_TT}, a_sgemm_null},
    {"irr_shortA_dgemm", 1, {a_blasdsp_irr_shortA_dgemm_NN, a_blasdsp_irr_shortA_dgemm_NT,
                             a_blasdsp_irr_shortA_dgemm_TN, a_blasdsp_irr_shortA_dgemm_TT}, a_dgemm_null},
    {"irr_shortA_sgemm", 0, {a_blasdsp_irr_shortA_sgemm_NN, a_blasdsp_irr_shortA_sgemm_NT,
                             a_blasdsp_irr_shortA_sgemm_TN, a_blasdsp_irr_shortA_sgemm_TT}, a_sgemm_null},
    {"irr_thinB_dgemm", 1, {a_blasdsp_irr_thinB_dgemm_NN, a_blasdsp_irr_thinB_dgemm_NT,
                            a_blasdsp_irr_thinB_dgemm_TN, a_blasdsp_irr_thinB_dgemm_TT}, a_dgemm_null},
    {"irr_thinB_sgemm", 0, {a_blasdsp_irr_thinB_sgemm_NN, a_blasdsp_irr_thinB_sgemm_NT,
                            a_blasdsp_irr_thinB_sgemm_TN, a_blasdsp_irr_thinB_sgemm_TT}, a_sgemm_null},
};

static void fill(int f64, void *X, int64_t rows, int64_t cols, int64_t ld, int trans, uint64_t seed)
{
    if (f64) BLAS_OK(blasdsp_fill_f64((double *)X, rows, cols, ld, trans, 0, seed, 0.0, NULL));
    else BLAS_OK(blasdsp_fill_f32((float *)X, rows, cols, ld, trans, 0, seed, 0.0f, NULL));
}

int main(int argc, char **argv)
{
    if (argc < 6) {
        printf("Usage: %s <family> dev_id M K N [num_threads]\n"
               "family: dgemm sgemm irr_shortA_dgemm irr_shortA_sgemm irr_thinB_dgemm irr_thinB_sgemm\n",
               argv[0]);
        return -1;
    }
    const struct family *fam = NULL;
    for (size_t i = 0; i < sizeof FAMILIES / sizeof FAMILIES[0]; ++i)
        if (strcmp(argv[1], FAMILIES[i].name) == 0) fam = &FAMILIES[i];
    if (!fam) {
        fprintf(stderr, "unknown family %s\n", argv[1]);
        return -1;
    }
    const int dev_id = atoi(argv[2]);
    const int64_t M = atol(argv[3]), K = atol(argv[4]), N = atol(argv[5]);
    if (M <= 0 || K <= 0 || N <= 0) {
        fprintf(stderr, "M, K, N must be positive\n");
        return -1;
    }
    HIP_OK(hipSetDevice(dev_id));
    printf("%s (%s) M=%ld, K=%ld, N=%ld\n", fam->name, blasdsp_version(), (long)M, (long)K, (long)N);

    /* the reference's padding (dgemm_demo.c:28-48) */
    const int64_t M0 = M + 64, K0 = K + 128, N0 = N + 256;
    const size_t s = fam->f64 ? 8 : 4;
    void *A, *AT, *B, *BT, *C[4];
    HIP_OK(hipMalloc(&A, s * M0 * K0));
    HIP_OK(hipMalloc(&AT, s * K0 * M0));
    HIP_OK(hipMalloc(&B, s * K0 * N0));
    HIP_OK(hipMalloc(&BT, s * N0 * K0));
    for (int m = 0; m < 4; ++m) HIP_OK(hipMalloc(&C[m], s * M0 * N0));
    fill(fam->f64, A, M, K, K0, 0, 1);
    fill(fam->f64, AT, M, K, M0, 1, 1);
    fill(fam->f64, B, K, N, N0, 0, 2);
    fill(fam->f64, BT, K, N, K0, 1, 2);
    HIP_OK(hipDeviceSynchronize());

    /* mode m: A or A^T (lda K0 / M0), B or B^T (ldb N0 / K0); ldc N0 */
    const void *a_of[4] = {A, A, AT, AT}, *b_of[4] = {B, BT, B, BT};
    const int64_t lda_of[4] = {K0, K0, M0, M0}, ldb_of[4] = {N0, K0, N0, K0};
    const char *mname[4] = {"NN", "NT", "TN", "TT"};

    /* launch overhead: the empty kernel, after one untimed launch (the first
     * launch of a process also loads the code object) */
    BLAS_OK(fam->null_fn(K0, N0, N0, M, N, K, A, B, C[0], 1.0, 1.0, NULL));
    HIP_OK(hipDeviceSynchronize());
    double t0 = timer_us();
    for (int i = 0; i < NUM_TESTS; ++i) {
        BLAS_OK(fam->null_fn(K0, N0, N0, M, N, K, A, B, C[0], 1.0, 1.0, NULL));
        HIP_OK(hipDeviceSynchronize());
    }
    const double t_ov = (timer_us() - t0) / NUM_TESTS;

    const double gflop = 2.0 * (double)M * (double)N * (double)K * 1e-9;
    for (int m = 0; m < 4; ++m) {
        /* warm-up call, then NUM_TESTS timed calls from C = 0 */
        BLAS_OK(fam->mode[m](lda_of[m], ldb_of[m], N0, M, N, K, a_of[m], b_of[m], C[m], 1.0, 0.0, NULL));
        HIP_OK(hipDeviceSynchronize());
        t0 = timer_us();
        for (int i = 0; i < NUM_TESTS; ++i)
            BLAS_OK(fam->mode[m](lda_of[m], ldb_of[m], N0, M, N, K, a_of[m], b_of[m], C[m], 1.0, 0.0, NULL));
        HIP_OK(hipDeviceSynchronize());
        const double t = (timer_us() - t0) / NUM_TESTS - t_ov;
        printf("%s_%s elapsed time: %lf seconds, GFLOPS: %.1f\n", fam->name, mname[m], t * 1e-6,
               gflop / (t * 1e-6));
    }

    /* the four results must agree */
    const size_t bytes = s * (size_t)M0 * (size_t)N0;
    unsigned char *h[4];
    for (int m = 0; m < 4; ++m) {
        h[m] = (unsigned char *)malloc(bytes);
        if (!h[m]) {
            fprintf(stderr, "out of host memory\n");
            return 3;
        }
        HIP_OK(hipMemcpy(h[m], C[m], bytes, hipMemcpyDeviceToHost));
    }
    double worst = 0.0;
    int identical = 1;
    for (int m = 1; m < 4; ++m) {
        double num = 0.0, den = 0.0;
        for (int64_t i = 0; i < M; ++i)
            for (int64_t j = 0; j < N; ++j) {
                const size_t at = (size_t)i * N0 + j;
                const double r = fam->f64 ? ((double *)h[0])[at] : ((float *)h[0])[at];
                const double x = fam->f64 ? ((double *)h[m])[at] : ((float *)h[m])[at];
                num += (x - r) * (x - r);
                den += r * r;
                if (x != r) identical = 0;
            }
        const double d = den > 0.0 ? sqrt(num / den) : sqrt(num);
        if (d > worst || d != d) worst = d;
    }
    const int same = worst <= (fam->f64 ? 1e-12 : 1e-5);
    printf("%s (largest normwise difference from NN: %.3e%s)\n", same ? "Success!" : "Result Error!", worst,
           identical ? ", all four bit-identical" : "");
    for (int m = 0; m < 4; ++m) {
        free(h[m]);
        HIP_OK(hipFree(C[m]));
    }
    HIP_OK(hipFree(A));
    HIP_OK(hipFree(AT));
    HIP_OK(hipFree(B));
    HIP_OK(hipFree(BT));
    return same ? 0 : 1;
}
