// blasdsp_api.hip — the C ABI: argument checks, kernel selection, workspace.
//
// Replaces the reference's launcher shims (launch_dgemm_NN & co.,
// dgemm_release_V1.1/device_code/dgemm_launcher.c:10-28,
// sgemm_release_V1.3.2/device_code/sgemm_launcher.c:11-37) and the
// hthread_group_exec-by-name boundary (dgemm_release_V1.1/host_code/
// dgemm_demo.c:151): the host calls a C function, which validates, picks a
// kernel and enqueues it on the caller's HIP stream.  Unlike the reference
// (void kernels, silent wrong results on unsupported shapes,
// dgemm_release_V1.1/readme.txt:11-15) every shape is supported and every
// call returns a status.
#include <errno.h>
#include <stdlib.h>

#include <map>
#include <mutex>
#include <utility>

#include "../../include/blasdsp.h"
#include "api_common.h"
#include "tune.h"

namespace blasdsp {

thread_local int g_last_hip_error = 0;

int hip_status(hipError_t e)
{
    if (e == hipSuccess) return BLASDSP_SUCCESS;
    g_last_hip_error = (int)e;
    if (e == hipErrorOutOfMemory) return BLASDSP_ERR_NOMEM;
    if (e == hipErrorInvalidValue) return BLASDSP_ERR_INVALID;
    return BLASDSP_ERR_DEVICE;
}

// Split-K workspace, one grow-only buffer per (device, stream): calls on one
// stream are ordered, so reuse within a stream is safe; distinct streams never
// share a buffer.  Growing frees the old buffer with hipFree, which waits for
// the device, so in-flight users of the old buffer finish first.
namespace {
struct Workspace {
    void *ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_ws_mutex;
std::map<std::pair<int, hipStream_t>, Workspace> g_ws;
}  // namespace

hipError_t workspace(hipStream_t s, size_t bytes, void **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    Workspace &w = g_ws[std::make_pair(dev, s)];
    if (w.bytes < bytes) {
        if (w.ptr) {
            e = hipFree(w.ptr);
            w.ptr = nullptr;
            w.bytes = 0;
            if (e != hipSuccess) return e;
        }
        size_t want = bytes + bytes / 4;  // headroom against creeping growth
        e = hipMalloc(&w.ptr, want);
        if (e != hipSuccess) {
            w.ptr = nullptr;
            return e;
        }
        w.bytes = want;
    }
    *out = w.ptr;
    return hipSuccess;
}

hipError_t workspace_release(hipStream_t s)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    auto it = g_ws.find(std::make_pair(dev, s));
    if (it == g_ws.end()) return hipSuccess;
    if (it->second.ptr) e = hipFree(it->second.ptr);
    g_ws.erase(it);
    return e;
}

int num_cus()
{
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cached[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
        cached[dev] = n;
    }
    return cached[dev];
}

bool parse_trans(char t, bool &out)
{
    switch (t) {
    case 'N': case 'n': out = false; return true;
    case 'T': case 't': case 'C': case 'c': out = true; return true;
    default: return false;
    }
}

namespace {

// The heuristic's skinny test: small side <= 48, a long side at least twice
// the small one, and enough panels and K slices to fill the GPU.  From 49
// rows both kernels pad the small side to 64, and there the tile kernel ran
// as fast or faster on MI355X: over 18 shapes of 64 rows (long side 3024 /
// 8192 / 20480, K 2048 / 8192 / 20480, FP64 and FP32) at most 2 % slower
// than the skinny kernel and up to 1.57x faster (FP32 8192 x 64 x 8192).
// Explicit BLASDSP_ALGO_SKINNY still takes up to 64 rows.
template <typename T>
bool skinny_shape(const Gemm<T> &g)
{
    const int64_t small = g.M <= g.N ? g.M : g.N;
    const int64_t large = g.M <= g.N ? g.N : g.M;
    return small <= 48 && large >= 2 * small && skinny_fits(g) && skinny_fills(g, num_cus());
}

}  // namespace

template <typename T>
int choose_kernel(const Gemm<T> &g, int algo, int splitk, KernelChoice &c)
{
    // 64 x 64 tiles when 128 x 128 tiles would leave most CUs idle
    const int64_t t128 = ((g.M + 127) / 128) * ((g.N + 127) / 128);
    c.tile = t128 < num_cus() ? 64 : 128;
    if (algo == BLASDSP_ALGO_TILE) c.skinny = false;
    else if (algo == BLASDSP_ALGO_SKINNY) c.skinny = skinny_fits(g);  // else: tile kernel
    else if (algo == BLASDSP_ALGO_AUTO) {
        TuneChoice t;
        if (tune_lookup(sizeof(T) == 8, g.ta, g.tb, g.M, g.N, g.K, &t) &&
            (t.algo == BLASDSP_ALGO_TILE || skinny_fits(g))) {
            c.skinny = t.algo == BLASDSP_ALGO_SKINNY;
            c.tile = t.tile;
            if (splitk <= 0) splitk = t.splitk;
        } else {
            c.skinny = skinny_shape(g);
        }
    } else {
        return BLASDSP_ERR_INVALID;
    }
    if (splitk <= 0) splitk = c.skinny ? skinny_pick_splitk(g, num_cus()) : tile_pick_splitk(g, c.tile, num_cus());
    if (splitk > 256) return BLASDSP_ERR_INVALID;
    c.splitk = splitk;
    return BLASDSP_SUCCESS;
}

template <typename T>
size_t kernel_scratch_bytes(const Gemm<T> &g, int splitk)
{
    return splitk > 1 ? (size_t)splitk * (size_t)g.M * (size_t)g.N * sizeof(T) : 0;
}

template <typename T>
int launch_kernel(const Gemm<T> &g, const KernelChoice &c, T *work, hipStream_t s)
{
    if (c.skinny) return hip_status(launch_skinny(g, c.splitk, work, s));
    return hip_status(launch_tile(g, c.tile, c.splitk, work, s));
}

template int choose_kernel<double>(const Gemm<double> &, int, int, KernelChoice &);
template int choose_kernel<float>(const Gemm<float> &, int, int, KernelChoice &);
template size_t kernel_scratch_bytes<double>(const Gemm<double> &, int);
template size_t kernel_scratch_bytes<float>(const Gemm<float> &, int);
template int launch_kernel<double>(const Gemm<double> &, const KernelChoice &, double *, hipStream_t);
template int launch_kernel<float>(const Gemm<float> &, const KernelChoice &, float *, hipStream_t);

namespace {

template <typename T>
int run(const Gemm<T> &g, int algo, int splitk, hipStream_t s)
{
    if (g.M == 0 || g.N == 0) return BLASDSP_SUCCESS;
    if (g.K == 0 || g.alpha == T(0)) return hip_status(launch_scale(g, s));
    KernelChoice c;
    int st = choose_kernel(g, algo, splitk, c);
    if (st != BLASDSP_SUCCESS) return st;
    T *work = nullptr;
    const size_t bytes = kernel_scratch_bytes(g, c.splitk);
    if (bytes > 0) {
        void *p = nullptr;
        hipError_t e = workspace(s, bytes, &p);
        if (e != hipSuccess) return hip_status(e);
        work = (T *)p;
    }
    return launch_kernel(g, c, work, s);
}

}  // namespace

template <typename T>
int gemm_entry(char ta, char tb, int64_t M, int64_t N, int64_t K, T alpha, const T *A, int64_t lda,
               const T *B, int64_t ldb, T beta, T *C, int64_t ldc, int algo, int splitk, void *stream)
{
    Gemm<T> g;
    int st = check<T>(g, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
    if (st != BLASDSP_SUCCESS) return st;
    return run(g, algo, splitk, (hipStream_t)stream);
}

template int gemm_entry<double>(char, char, int64_t, int64_t, int64_t, double, const double *, int64_t,
                                const double *, int64_t, double, double *, int64_t, int, int, void *);
template int gemm_entry<float>(char, char, int64_t, int64_t, int64_t, float, const float *, int64_t,
                               const float *, int64_t, float, float *, int64_t, int, int, void *);

template <typename T>
int null_entry(char ta, char tb, int64_t lda, int64_t ldb, int64_t ldc, int64_t M, int64_t N,
               int64_t K, const T *A, const T *B, T *C, T alpha, T beta, void *stream)
{
    Gemm<T> g;
    int st = check<T>(g, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
    if (st != BLASDSP_SUCCESS) return st;
    return hip_status(launch_null((hipStream_t)stream));
}

template int null_entry<double>(char, char, int64_t, int64_t, int64_t, int64_t, int64_t, int64_t, const double *,
                                const double *, double *, double, double, void *);
template int null_entry<float>(char, char, int64_t, int64_t, int64_t, int64_t, int64_t, int64_t, const float *,
                               const float *, float *, float, float, void *);

}  // namespace blasdsp

using blasdsp::gemm_entry;

extern "C" {

const char *blasdsp_version(void) { return "blasdsp 0.1.0 (gfx950)"; }

const char *blasdsp_status_string(int status)
{
    switch (status) {
    case BLASDSP_SUCCESS: return "success";
    case BLASDSP_ERR_INVALID: return "invalid argument";
    case BLASDSP_ERR_NOMEM: return "out of memory";
    case BLASDSP_ERR_DEVICE: return "HIP runtime error";
    case BLASDSP_ERR_UNSUPPORTED: return "unsupported";
    default: return "unknown status";
    }
}

int blasdsp_last_hip_error(void) { return blasdsp::g_last_hip_error; }

int blasdsp_dgemm(char transA, char transB, int64_t M, int64_t N, int64_t K, double alpha,
                  const double *A, int64_t lda, const double *B, int64_t ldb, double beta, double *C,
                  int64_t ldc, void *stream)
{
    return gemm_entry<double>(transA, transB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                              BLASDSP_ALGO_AUTO, 0, stream);
}

int blasdsp_sgemm(char transA, char transB, int64_t M, int64_t N, int64_t K, float alpha,
                  const float *A, int64_t lda, const float *B, int64_t ldb, float beta, float *C,
                  int64_t ldc, void *stream)
{
    return gemm_entry<float>(transA, transB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                             BLASDSP_ALGO_AUTO, 0, stream);
}

int blasdsp_dgemm_ex(char transA, char transB, int64_t M, int64_t N, int64_t K, double alpha,
                     const double *A, int64_t lda, const double *B, int64_t ldb, double beta,
                     double *C, int64_t ldc, int algo, int splitk, void *stream)
{
    return gemm_entry<double>(transA, transB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, algo,
                              splitk, stream);
}

int blasdsp_sgemm_ex(char transA, char transB, int64_t M, int64_t N, int64_t K, float alpha,
                     const float *A, int64_t lda, const float *B, int64_t ldb, float beta, float *C,
                     int64_t ldc, int algo, int splitk, void *stream)
{
    return gemm_entry<float>(transA, transB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, algo,
                             splitk, stream);
}

#define BLASDSP_MODE_ENTRY(name, T, ta, tb, algo)                                                 \
    int name(int64_t lda, int64_t ldb, int64_t ldc, int64_t M, int64_t N, int64_t K, const T *A,  \
             const T *B, T *C, T alpha, T beta, void *stream)                                    \
    {                                                                                             \
        return gemm_entry<T>(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, algo, 0,       \
                             stream);                                                             \
    }

BLASDSP_MODE_ENTRY(dgemm_NN, double, 'N', 'N', BLASDSP_ALGO_AUTO)
BLASDSP_MODE_ENTRY(dgemm_NT, double, 'N', 'T', BLASDSP_ALGO_AUTO)
BLASDSP_MODE_ENTRY(dgemm_TN, double, 'T', 'N', BLASDSP_ALGO_AUTO)
BLASDSP_MODE_ENTRY(dgemm_TT, double, 'T', 'T', BLASDSP_ALGO_AUTO)
BLASDSP_MODE_ENTRY(blasdsp_sgemm_NN, float, 'N', 'N', BLASDSP_ALGO_AUTO)
BLASDSP_MODE_ENTRY(blasdsp_sgemm_NT, float, 'N', 'T', BLASDSP_ALGO_AUTO)
BLASDSP_MODE_ENTRY(blasdsp_sgemm_TN, float, 'T', 'N', BLASDSP_ALGO_AUTO)
BLASDSP_MODE_ENTRY(blasdsp_sgemm_TT, float, 'T', 'T', BLASDSP_ALGO_AUTO)
BLASDSP_MODE_ENTRY(blasdsp_irr_shortA_dgemm_NN, double, 'N', 'N', BLASDSP_ALGO_SKINNY)
BLASDSP_MODE_ENTRY(blasdsp_irr_shortA_dgemm_NT, double, 'N', 'T', BLASDSP_ALGO_SKINNY)
BLASDSP_MODE_ENTRY(blasdsp_irr_shortA_dgemm_TN, double, 'T', 'N', BLASDSP_ALGO_SKINNY)
BLASDSP_MODE_ENTRY(blasdsp_irr_shortA_dgemm_TT, double, 'T', 'T', BLASDSP_ALGO_SKINNY)
BLASDSP_MODE_ENTRY(blasdsp_irr_shortA_sgemm_NN, float, 'N', 'N', BLASDSP_ALGO_SKINNY)
BLASDSP_MODE_ENTRY(blasdsp_irr_shortA_sgemm_NT, float, 'N', 'T', BLASDSP_ALGO_SKINNY)
BLASDSP_MODE_ENTRY(blasdsp_irr_shortA_sgemm_TN, float, 'T', 'N', BLASDSP_ALGO_SKINNY)
BLASDSP_MODE_ENTRY(blasdsp_irr_shortA_sgemm_TT, float, 'T', 'T', BLASDSP_ALGO_SKINNY)
BLASDSP_MODE_ENTRY(blasdsp_irr_thinB_dgemm_NN, double, 'N', 'N', BLASDSP_ALGO_SKINNY)
BLASDSP_MODE_ENTRY(blasdsp_irr_thinB_dgemm_NT, double, 'N', 'T', BLASDSP_ALGO_SKINNY)
BLASDSP_MODE_ENTRY(blasdsp_irr_thinB_dgemm_TN, double, 'T', 'N', BLASDSP_ALGO_SKINNY)
BLASDSP_MODE_ENTRY(blasdsp_irr_thinB_dgemm_TT, double, 'T', 'T', BLASDSP_ALGO_SKINNY)
BLASDSP_MODE_ENTRY(blasdsp_irr_thinB_sgemm_NN, float, 'N', 'N', BLASDSP_ALGO_SKINNY)
BLASDSP_MODE_ENTRY(blasdsp_irr_thinB_sgemm_NT, float, 'N', 'T', BLASDSP_ALGO_SKINNY)
BLASDSP_MODE_ENTRY(blasdsp_irr_thinB_sgemm_TN, float, 'T', 'N', BLASDSP_ALGO_SKINNY)
BLASDSP_MODE_ENTRY(blasdsp_irr_thinB_sgemm_TT, float, 'T', 'T', BLASDSP_ALGO_SKINNY)
#undef BLASDSP_MODE_ENTRY

int dgemm_null(int64_t lda, int64_t ldb, int64_t ldc, int64_t M, int64_t N, int64_t K,
               const double *A, const double *B, double *C, double alpha, double beta, void *stream)
{
    return blasdsp::null_entry<double>('N', 'N', lda, ldb, ldc, M, N, K, A, B, C, alpha, beta, stream);
}


}  // extern "C"

namespace blasdsp {
namespace {
constexpr int kSchemeFP8 = 100;  // split_entry: the scaled FP8 form rather than a split scheme

int split_entry(char transA, char transB, int64_t M, int64_t N, int64_t K, float alpha, const float *A,
                int64_t lda, const float *B, int64_t ldb, float beta, float *C, int64_t ldc, int scheme,
                void *stream)
{
    Gemm<float> g;
    int st = check<float>(g, transA, transB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
    if (st != BLASDSP_SUCCESS) return st;
    if (M == 0 || N == 0) return BLASDSP_SUCCESS;
    hipStream_t s = (hipStream_t)stream;
    if (K == 0 || alpha == 0.0f) return hip_status(launch_scale(g, s));
    void *work = nullptr;
    const size_t bytes = scheme == kSchemeFP8 ? fp8_workspace_bytes(M, N, K) : split_workspace_bytes(M, N, K, scheme);
    hipError_t e = workspace(s, bytes, &work);
    if (e != hipSuccess) return hip_status(e);
    if (scheme == kSchemeFP8) return hip_status(launch_fp8_gemm(g, work, s));
    return hip_status(launch_split_gemm(g, scheme, work, s));
}
}  // namespace
}  // namespace blasdsp

extern "C" {

int blasdsp_sgemm_split_bf16(char transA, char transB, int64_t M, int64_t N, int64_t K, float alpha,
                             const float *A, int64_t lda, const float *B, int64_t ldb, float beta,
                             float *C, int64_t ldc, int nterms, void *stream)
{
    if (nterms != 3 && nterms != 6) return BLASDSP_ERR_INVALID;
    return blasdsp::split_entry(transA, transB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                                nterms == 6 ? 0 : 1, stream);
}

int blasdsp_sgemm_split_fp16(char transA, char transB, int64_t M, int64_t N, int64_t K, float alpha,
                             const float *A, int64_t lda, const float *B, int64_t ldb, float beta,
                             float *C, int64_t ldc, void *stream)
{
    return blasdsp::split_entry(transA, transB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, 2, stream);
}

int blasdsp_sgemm_fp8(char transA, char transB, int64_t M, int64_t N, int64_t K, float alpha, const float *A,
                      int64_t lda, const float *B, int64_t ldb, float beta, float *C, int64_t ldc, void *stream)
{
    return blasdsp::split_entry(transA, transB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, blasdsp::kSchemeFP8,
                                stream);
}

int blasdsp_stream_release(void *stream)
{
    const hipError_t e1 = blasdsp::workspace_release((hipStream_t)stream);
    const hipError_t e2 = blasdsp::grouped_plan_release((hipStream_t)stream);
    return blasdsp::hip_status(e1 != hipSuccess ? e1 : e2);
}

int blasdsp_fill_f64(double *X, int64_t rows, int64_t cols, int64_t ld, int trans, int dist,
                     uint64_t seed, double pad, void *stream)
{
    if (rows < 0 || cols < 0 || dist < 0 || dist > 5) return BLASDSP_ERR_INVALID;
    if (ld < ((trans ? rows : cols) > 1 ? (trans ? rows : cols) : 1)) return BLASDSP_ERR_INVALID;
    if (rows > 0 && cols > 0 && !X) return BLASDSP_ERR_INVALID;
    return blasdsp::hip_status(
        blasdsp::launch_fill<double>(X, rows, cols, ld, trans, dist, seed, pad, (hipStream_t)stream));
}

int blasdsp_fill_f32(float *X, int64_t rows, int64_t cols, int64_t ld, int trans, int dist,
                     uint64_t seed, float pad, void *stream)
{
    if (rows < 0 || cols < 0 || dist < 0 || dist > 5) return BLASDSP_ERR_INVALID;
    if (ld < ((trans ? rows : cols) > 1 ? (trans ? rows : cols) : 1)) return BLASDSP_ERR_INVALID;
    if (rows > 0 && cols > 0 && !X) return BLASDSP_ERR_INVALID;
    return blasdsp::hip_status(
        blasdsp::launch_fill<float>(X, rows, cols, ld, trans, dist, seed, pad, (hipStream_t)stream));
}

}  // extern "C"
