// ref_api.hip — the reference's own call surface: its kernel names, argument
// order and semantics, with the barrier ids dropped and a stream appended.
//
// The reference declares (all operands packed: A is M x K, or K x M for the T
// modes; B is K x N, or N x K; C is M x N):
//   sgemm_{NN,NT,TT}_qxx(M, N, K, b_id, A, B, C, alpha, beta),
//   sgemm_TN_qxx(M, N, K, bid_main, bid_T, A, B, C, alpha, beta)
//       C = alpha op(A) op(B) + beta, beta a SCALAR, input C ignored
//       (sgemm_release_V1.3.2/include/sgemm_kernel.h:11-25;
//        semantics sgemm_release_V1.3.2/device_code/sgemm_kernel.c:2681,
//        2745-2749)
//   irr_shortA_dgemm_{NN,NT,TN,TT}_dbuf(M, N, K, b_id[, bid_T], A, B, C)
//       C += op(A) op(B)   (irr_shortA_dgemm_release_V0.0/include/
//        irr_shortA_dgemm_kernel.h:11-25)
//   irr_shortA_sgemm_{NN,NT,TN,TT}(...)          same, FP32
//       (irr_shortA_sgemm_release_V0.0/include/irr_shortA_sgemm_kernel.h:11-25)
//   irr_thinB_dgemm_{NN,NT,TN,TT}_dbuf(...)      same, small N
//       (irr_thinB_dgemm_release_V0.0/include/irr_thinB_dgemm_kernel.h:11-25)
//   irr_thinB_sgemm_{NN,NT,TN,TT}(M, N, K, b_id[, bid_T], A, B, C, alpha, beta)
//       C = alpha op(A) op(B) + beta (scalar), input C ignored
//       (irr_thinB_sgemm_release_V1.2/include/irr_thinB_sgemm_kernel.h:11-25;
//        semantics irr_thinB_sgemm_kernel.c:3095-3112, 3182-3193)
//   *_null: empty kernels of the same argument lists.
// The scalar-beta forms write C = beta first (one pass over C, the role of
// the reference's C_plus_beta, sgemm_kernel.c:25-160) and then run the BLAS
// GEMM with beta = 1: C = alpha op(A) op(B) + C.  The irr names take the
// bandwidth-bound skinny kernel (the tile kernel where it does not fit:
// small side above 64 or per-lane offsets past 2^31 bytes).
//
// blasdsp_group_exec replaces the reference host's by-name launch,
// hthread_group_exec(tg, "launch_dgemm_NN", n_scalar, n_ptr, args)
// (dgemm_release_V1.1/host_code/dgemm_demo.c:151): the same launcher names
// and the same 64-bit argument words (scalars first, then pointers), with the
// launchers' compiled-in alpha = 5, beta = 3 for the sgemm and irr_thinB_sgemm
// families (sgemm_release_V1.3.2/device_code/sgemm_launcher.c:11-12,
// irr_thinB_sgemm_release_V1.2/device_code/irr_thinB_sgemm_launcher.c:10-11).
#include <string.h>

#include "api_common.h"

namespace blasdsp {
namespace {

template <typename T>
__global__ void set_kernel(T *C, int64_t total, T v)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x)
        C[e] = v;
}

// Packed leading dimensions of the reference's operands (at least 1, the
// leading dimension of an empty matrix).
struct Packed {
    int64_t lda, ldb, ldc;
};
Packed packed(char ta, char tb, int64_t M, int64_t N, int64_t K)
{
    auto ld = [](int64_t cols) { return cols > 1 ? cols : (int64_t)1; };
    return Packed{ld(ta == 'T' ? M : K), ld(tb == 'T' ? K : N), ld(N)};
}

// C += op(A) op(B) on packed operands (the dgemm / irr semantics).
template <typename T>
int accumulate(char ta, char tb, int64_t M, int64_t N, int64_t K, const T *A, const T *B, T *C, int algo,
               void *stream)
{
    const Packed p = packed(ta, tb, M, N, K);
    return gemm_entry<T>(ta, tb, M, N, K, T(1), A, p.lda, B, p.ldb, T(1), C, p.ldc, algo, 0, stream);
}

// C = alpha op(A) op(B) + beta (scalar) on packed operands; input C ignored.
template <typename T>
int scalar_beta(char ta, char tb, int64_t M, int64_t N, int64_t K, const T *A, const T *B, T *C, T alpha,
                T beta, int algo, void *stream)
{
    const Packed p = packed(ta, tb, M, N, K);
    Gemm<T> g;
    int st = check<T>(g, ta, tb, M, N, K, alpha, A, p.lda, B, p.ldb, T(1), C, p.ldc);
    if (st != BLASDSP_SUCCESS) return st;
    if (M == 0 || N == 0) return BLASDSP_SUCCESS;
    const int64_t total = M * N;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL((set_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, C, total,
                       beta);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_status(e);
    if (K == 0 || alpha == T(0)) return BLASDSP_SUCCESS;
    return gemm_entry<T>(ta, tb, M, N, K, alpha, A, p.lda, B, p.ldb, T(1), C, p.ldc, algo, 0, stream);
}

template <typename T>
int null_packed(int64_t M, int64_t N, int64_t K, const T *A, const T *B, T *C, T alpha, T beta, void *stream)
{
    const Packed p = packed('N', 'N', M, N, K);
    return null_entry<T>('N', 'N', p.lda, p.ldb, p.ldc, M, N, K, A, B, C, alpha, beta, stream);
}

}  // namespace
}  // namespace blasdsp

using blasdsp::accumulate;
using blasdsp::null_packed;
using blasdsp::scalar_beta;

extern "C" {

#define BLASDSP_QXX(mode, ta, tb)                                                                         \
    int sgemm_##mode##_qxx(int64_t M, int64_t N, int64_t K, const float *A, const float *B, float *C,        \
                           float alpha, float beta, void *stream)                                        \
    {                                                                                                     \
        return scalar_beta<float>(ta, tb, M, N, K, A, B, C, alpha, beta, BLASDSP_ALGO_AUTO, stream);      \
    }
BLASDSP_QXX(NN, 'N', 'N')
BLASDSP_QXX(NT, 'N', 'T')
BLASDSP_QXX(TN, 'T', 'N')
BLASDSP_QXX(TT, 'T', 'T')
#undef BLASDSP_QXX

int sgemm_null(int64_t M, int64_t N, int64_t K, const float *A, const float *B, float *C, float alpha, float beta,
               void *stream)
{
    return null_packed<float>(M, N, K, A, B, C, alpha, beta, stream);
}

#define BLASDSP_IRR_ACC(name, T, ta, tb)                                                                  \
    int name(int64_t M, int64_t N, int64_t K, const T *A, const T *B, T *C, void *stream)                 \
    {                                                                                                     \
        return accumulate<T>(ta, tb, M, N, K, A, B, C, BLASDSP_ALGO_SKINNY, stream);                      \
    }
#define BLASDSP_IRR_NULL(name, T)                                                                         \
    int name(int64_t M, int64_t N, int64_t K, const T *A, const T *B, T *C, void *stream)                 \
    {                                                                                                     \
        return null_packed<T>(M, N, K, A, B, C, T(1), T(1), stream);                                      \
    }
BLASDSP_IRR_ACC(irr_shortA_dgemm_NN_dbuf, double, 'N', 'N')
BLASDSP_IRR_ACC(irr_shortA_dgemm_NT_dbuf, double, 'N', 'T')
BLASDSP_IRR_ACC(irr_shortA_dgemm_TN_dbuf, double, 'T', 'N')
BLASDSP_IRR_ACC(irr_shortA_dgemm_TT_dbuf, double, 'T', 'T')
BLASDSP_IRR_NULL(irr_shortA_dgemm_null, double)
BLASDSP_IRR_ACC(irr_shortA_sgemm_NN, float, 'N', 'N')
BLASDSP_IRR_ACC(irr_shortA_sgemm_NT, float, 'N', 'T')
BLASDSP_IRR_ACC(irr_shortA_sgemm_TN, float, 'T', 'N')
BLASDSP_IRR_ACC(irr_shortA_sgemm_TT, float, 'T', 'T')
BLASDSP_IRR_NULL(irr_shortA_sgemm_null, float)
BLASDSP_IRR_ACC(irr_thinB_dgemm_NN_dbuf, double, 'N', 'N')
BLASDSP_IRR_ACC(irr_thinB_dgemm_NT_dbuf, double, 'N', 'T')
BLASDSP_IRR_ACC(irr_thinB_dgemm_TN_dbuf, double, 'T', 'N')
BLASDSP_IRR_ACC(irr_thinB_dgemm_TT_dbuf, double, 'T', 'T')
BLASDSP_IRR_NULL(irr_thinB_dgemm_null, double)
#undef BLASDSP_IRR_ACC
#undef BLASDSP_IRR_NULL

#define BLASDSP_THINB_S(mode, ta, tb)                                                                     \
    int irr_thinB_sgemm_##mode(int64_t M, int64_t N, int64_t K, const float *A, const float *B, float *C,   \
                               float alpha, float beta, void *stream)                                    \
    {                                                                                                     \
        return scalar_beta<float>(ta, tb, M, N, K, A, B, C, alpha, beta, BLASDSP_ALGO_SKINNY, stream);    \
    }
BLASDSP_THINB_S(NN, 'N', 'N')
BLASDSP_THINB_S(NT, 'N', 'T')
BLASDSP_THINB_S(TN, 'T', 'N')
BLASDSP_THINB_S(TT, 'T', 'T')
#undef BLASDSP_THINB_S

int irr_thinB_sgemm_null(int64_t M, int64_t N, int64_t K, const float *A, const float *B, float *C, float alpha,
                         float beta, void *stream)
{
    return null_packed<float>(M, N, K, A, B, C, alpha, beta, stream);
}

}  // extern "C"

// ---- by-name launch ---------------------------------------------------------

namespace blasdsp {
namespace {

enum Family { kDgemm, kSgemm, kShortAD, kShortAS, kThinBD, kThinBS };

struct Launcher {
    const char *name;
    Family fam;
    char ta, tb;      // 0: the family's empty kernel
    int n_scalar;     // the reference launcher's scalar words (dims + barrier ids)
};

// Launcher names and scalar-word counts of the reference's launcher files
// (dgemm_release_V1.1/device_code/dgemm_launcher.c:10-26, sgemm_release_V1.3.2/
// device_code/sgemm_launcher.c:14-35, irr_*_launcher.c:10-30): one barrier id,
// two for the modes that use transposer cores.
const Launcher kLaunchers[] = {
    {"launch_dgemm_NN", kDgemm, 'N', 'N', 7},  {"launch_dgemm_NT", kDgemm, 'N', 'T', 7},
    {"launch_dgemm_TN", kDgemm, 'T', 'N', 8},  {"launch_dgemm_TT", kDgemm, 'T', 'T', 7},
    {"launch_dgemm_null", kDgemm, 0, 0, 7},
    {"launch_sgemm_NN", kSgemm, 'N', 'N', 4},  {"launch_sgemm_NT", kSgemm, 'N', 'T', 4},
    {"launch_sgemm_TN", kSgemm, 'T', 'N', 5},  {"launch_sgemm_TT", kSgemm, 'T', 'T', 4},
    {"launch_sgemm_null", kSgemm, 0, 0, 4},
    {"launch_irr_shortA_dgemm_NN", kShortAD, 'N', 'N', 4}, {"launch_irr_shortA_dgemm_NT", kShortAD, 'N', 'T', 4},
    {"launch_irr_shortA_dgemm_TN", kShortAD, 'T', 'N', 5}, {"launch_irr_shortA_dgemm_TT", kShortAD, 'T', 'T', 5},
    {"launch_irr_shortA_dgemm_null", kShortAD, 0, 0, 4},
    {"launch_irr_shortA_sgemm_NN", kShortAS, 'N', 'N', 4}, {"launch_irr_shortA_sgemm_NT", kShortAS, 'N', 'T', 4},
    {"launch_irr_shortA_sgemm_TN", kShortAS, 'T', 'N', 5}, {"launch_irr_shortA_sgemm_TT", kShortAS, 'T', 'T', 5},
    {"launch_irr_shortA_sgemm_null", kShortAS, 0, 0, 4},
    {"launch_irr_thinB_dgemm_NN", kThinBD, 'N', 'N', 5},   {"launch_irr_thinB_dgemm_NT", kThinBD, 'N', 'T', 4},
    {"launch_irr_thinB_dgemm_TN", kThinBD, 'T', 'N', 5},   {"launch_irr_thinB_dgemm_TT", kThinBD, 'T', 'T', 4},
    {"launch_irr_thinB_dgemm_null", kThinBD, 0, 0, 4},
    {"launch_irr_thinB_sgemm_NN", kThinBS, 'N', 'N', 5},   {"launch_irr_thinB_sgemm_NT", kThinBS, 'N', 'T', 4},
    {"launch_irr_thinB_sgemm_TN", kThinBS, 'T', 'N', 5},   {"launch_irr_thinB_sgemm_TT", kThinBS, 'T', 'T', 4},
    {"launch_irr_thinB_sgemm_null", kThinBS, 0, 0, 4},
};

constexpr float kLaunchAlpha = 5.0f, kLaunchBeta = 3.0f;

int exec(const Launcher &L, const uint64_t *a, void *stream)
{
    const int s = L.n_scalar;
    if (L.fam == kDgemm) {
        // lda, ldb, ldc, M, N, K, barrier id(s), A, B, C;  C += op(A) op(B)
        const int64_t lda = (int64_t)a[0], ldb = (int64_t)a[1], ldc = (int64_t)a[2];
        const int64_t M = (int64_t)a[3], N = (int64_t)a[4], K = (int64_t)a[5];
        const double *A = (const double *)a[s], *B = (const double *)a[s + 1];
        double *C = (double *)a[s + 2];
        if (!L.ta) return null_entry<double>('N', 'N', lda, ldb, ldc, M, N, K, A, B, C, 1.0, 1.0, stream);
        return gemm_entry<double>(L.ta, L.tb, M, N, K, 1.0, A, lda, B, ldb, 1.0, C, ldc, BLASDSP_ALGO_AUTO, 0,
                                  stream);
    }
    // M, N, K, barrier id(s), A, B, C (packed)
    const int64_t M = (int64_t)a[0], N = (int64_t)a[1], K = (int64_t)a[2];
    const bool f64 = L.fam == kShortAD || L.fam == kThinBD;
    const bool scal = L.fam == kSgemm || L.fam == kThinBS;
    const int algo = L.fam == kSgemm ? BLASDSP_ALGO_AUTO : BLASDSP_ALGO_SKINNY;
    if (f64) {
        const double *A = (const double *)a[s], *B = (const double *)a[s + 1];
        double *C = (double *)a[s + 2];
        if (!L.ta) return null_packed<double>(M, N, K, A, B, C, 1.0, 1.0, stream);
        return accumulate<double>(L.ta, L.tb, M, N, K, A, B, C, algo, stream);
    }
    const float *A = (const float *)a[s], *B = (const float *)a[s + 1];
    float *C = (float *)a[s + 2];
    if (!L.ta) return null_packed<float>(M, N, K, A, B, C, kLaunchAlpha, kLaunchBeta, stream);
    if (scal) return scalar_beta<float>(L.ta, L.tb, M, N, K, A, B, C, kLaunchAlpha, kLaunchBeta, algo, stream);
    return accumulate<float>(L.ta, L.tb, M, N, K, A, B, C, algo, stream);
}

}  // namespace
}  // namespace blasdsp

extern "C" int blasdsp_group_exec(const char *name, int n_scalar, int n_ptr, const uint64_t *args, void *stream)
{
    if (!name || !args) return BLASDSP_ERR_INVALID;
    for (const blasdsp::Launcher &L : blasdsp::kLaunchers) {
        if (strcmp(L.name, name) != 0) continue;
        if (n_scalar != L.n_scalar || n_ptr != 3) return BLASDSP_ERR_INVALID;
        for (int i = 0; i < (L.fam == blasdsp::kDgemm ? 6 : 3); ++i)
            if (args[i] > (uint64_t)INT64_MAX) return BLASDSP_ERR_INVALID;
        return blasdsp::exec(L, args, stream);
    }
    return BLASDSP_ERR_UNSUPPORTED;
}
