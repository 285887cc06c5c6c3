// tune.hip — the autotuner: per shape class, time the candidate kernels on
// the device and remember the fastest; the AUTO dispatch and the grouped
// path consult the table.
//
// No counterpart in the reference, whose blocking constants are compiled in
// per kernel (e.g. M_G = 576, K_G = 512, N_A = 64 at
// dgemm_release_V1.1/device_code/dgemm_kernel.c:2005-2011) and whose shape
// limits are documented rather than tuned (dgemm_release_V1.1/readme.txt:11-25).
//
// Shape class: precision, transpose modes and round(log2) of M, N and K.
// Candidates: the tile kernel with 64 and 128 tiles, each whole or split into
// 2, 4 or 8 K slices or into the tile cost model's slice count, and, when the
// smaller of M and N is at most 64, the skinny kernel with split-K factors
// 1 .. 64 (slices of at least 256).  Each candidate runs once to warm up and
// then three times between events on the caller's stream.  The table is a text file:
//   blasdsp-tune-v1 <dtype> <ta> <tb> <log2 M> <log2 N> <log2 K> <algo> <tile> <splitk> <microseconds>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "api_common.h"
#include "tune.h"

namespace blasdsp {
namespace {

typedef std::tuple<int, int, int, int, int, int> Key;
std::mutex g_tune_mutex;
std::map<Key, TuneChoice> g_table;

int bucket(int64_t x)
{
    if (x <= 1) return 0;
    int b = (int)lround(log2((double)x));
    return b < 0 ? 0 : (b > 40 ? 40 : b);
}

Key key_of(bool f64, bool ta, bool tb, int64_t M, int64_t N, int64_t K)
{
    return Key(f64 ? 1 : 0, ta ? 1 : 0, tb ? 1 : 0, bucket(M), bucket(N), bucket(K));
}

template <typename T>
hipError_t run_candidate(const Gemm<T> &g, const TuneChoice &c, T *work, hipStream_t s)
{
    if (c.algo == BLASDSP_ALGO_SKINNY) return launch_skinny(g, c.splitk, work, s);
    return launch_tile(g, c.tile, c.splitk, work, s);
}

template <typename T>
int autotune(bool ta, bool tb, int64_t M, int64_t N, int64_t K, hipStream_t s, TuneChoice *out)
{
    if (M <= 0 || N <= 0 || K <= 0) return BLASDSP_ERR_INVALID;
    std::vector<TuneChoice> cands;
    int max_split = 1;
    Gemm<T> shape{};
    shape.ta = ta; shape.tb = tb;
    shape.M = M; shape.N = N; shape.K = K;
    shape.lda = ta ? M : K; shape.ldb = tb ? K : N; shape.ldc = N;
    // tile kernel: 64 and 128 tiles, each whole or split into 2 .. 8 K slices
    // of at least 256, plus the slice count of the tile kernel's own cost
    // model when it is none of those
    for (int bm : {64, 128}) {
        for (int sk = 1; sk <= 8; sk *= 2) {
            if (sk > 1 && K / sk < 256) break;
            cands.push_back(TuneChoice{BLASDSP_ALGO_TILE, bm, sk, 0});
            max_split = sk > max_split ? sk : max_split;
        }
        const int model = tile_pick_splitk(shape, bm, num_cus());
        if (model > 1 && (model & (model - 1)) != 0) {
            cands.push_back(TuneChoice{BLASDSP_ALGO_TILE, bm, model, 0});
            max_split = model > max_split ? model : max_split;
        }
    }
    if (skinny_fits(shape)) {
        for (int sk = 1; sk <= 64; sk *= 2) {
            if (sk > 1 && K / sk < 256) break;
            cands.push_back(TuneChoice{BLASDSP_ALGO_SKINNY, 0, sk, 0});
            max_split = sk > max_split ? sk : max_split;
        }
    }
    // Dense operands in their stored orientation, U[-1,1).
    const int64_t ar = ta ? K : M, ac = ta ? M : K, br = tb ? N : K, bc = tb ? K : N;
    T *A = nullptr, *B = nullptr, *C = nullptr, *W = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&A, sizeof(T) * ar * ac);
    if (e == hipSuccess) e = hipMalloc(&B, sizeof(T) * br * bc);
    if (e == hipSuccess) e = hipMalloc(&C, sizeof(T) * M * N);
    if (e == hipSuccess && max_split > 1) e = hipMalloc(&W, sizeof(T) * max_split * M * N);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = launch_fill<T>(A, ar, ac, ac, 0, 0, 11, T(0), s);
    if (e == hipSuccess) e = launch_fill<T>(B, br, bc, bc, 0, 0, 12, T(0), s);
    if (e == hipSuccess) e = launch_fill<T>(C, M, N, N, 0, 0, 13, T(0), s);
    Gemm<T> g{ta, tb, M, N, K, T(1), T(0), A, ac, B, bc, C, N};
    TuneChoice best{BLASDSP_ALGO_TILE, 128, 1, 1e300};
    for (size_t i = 0; i < cands.size() && e == hipSuccess; ++i) {
        TuneChoice &c = cands[i];
        e = run_candidate(g, c, W, s);
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        for (int r = 0; r < 3 && e == hipSuccess; ++r) e = run_candidate(g, c, W, s);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        c.us = 1e3 * (double)ms / 3.0;
        if (e == hipSuccess && c.us < best.us) best = c;
    }
    if (e1) (void)hipEventDestroy(e1);
    if (e0) (void)hipEventDestroy(e0);
    // hipFree waits for the device, so the timed kernels have finished
    if (W) (void)hipFree(W);
    if (C) (void)hipFree(C);
    if (B) (void)hipFree(B);
    if (A) (void)hipFree(A);
    if (e != hipSuccess) return hip_status(e);
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        g_table[key_of(sizeof(T) == 8, ta, tb, M, N, K)] = best;
    }
    if (out) *out = best;
    return BLASDSP_SUCCESS;
}

}  // namespace

bool tune_lookup(bool f64, bool ta, bool tb, int64_t M, int64_t N, int64_t K, TuneChoice *out)
{
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    if (g_table.empty()) return false;
    auto it = g_table.find(key_of(f64, ta, tb, M, N, K));
    if (it == g_table.end()) return false;
    *out = it->second;
    return true;
}

}  // namespace blasdsp

extern "C" {

int blasdsp_autotune(int dtype, char transA, char transB, int64_t M, int64_t N, int64_t K, void *stream,
                     blasdsp_tune_choice *out)
{
    bool ta, tb;
    if (!blasdsp::parse_trans(transA, ta) || !blasdsp::parse_trans(transB, tb)) return BLASDSP_ERR_INVALID;
    blasdsp::TuneChoice c{};
    int st;
    if (dtype == BLASDSP_FLOAT64) st = blasdsp::autotune<double>(ta, tb, M, N, K, (hipStream_t)stream, &c);
    else if (dtype == BLASDSP_FLOAT32) st = blasdsp::autotune<float>(ta, tb, M, N, K, (hipStream_t)stream, &c);
    else return BLASDSP_ERR_INVALID;
    if (st == BLASDSP_SUCCESS && out) {
        out->algo = c.algo;
        out->tile = c.tile;
        out->splitk = c.splitk;
        out->time_us = c.us;
    }
    return st;
}

int blasdsp_tune_query(int dtype, char transA, char transB, int64_t M, int64_t N, int64_t K,
                       blasdsp_tune_choice *out)
{
    bool ta, tb;
    if (!out || !blasdsp::parse_trans(transA, ta) || !blasdsp::parse_trans(transB, tb)) return BLASDSP_ERR_INVALID;
    if (dtype != BLASDSP_FLOAT32 && dtype != BLASDSP_FLOAT64) return BLASDSP_ERR_INVALID;
    blasdsp::TuneChoice c;
    if (!blasdsp::tune_lookup(dtype == BLASDSP_FLOAT64, ta, tb, M, N, K, &c)) return BLASDSP_ERR_UNSUPPORTED;
    out->algo = c.algo;
    out->tile = c.tile;
    out->splitk = c.splitk;
    out->time_us = c.us;
    return BLASDSP_SUCCESS;
}

int blasdsp_tune_save(const char *path)
{
    if (!path) return BLASDSP_ERR_INVALID;
    FILE *f = fopen(path, "w");
    if (!f) return BLASDSP_ERR_INVALID;
    std::lock_guard<std::mutex> lock(blasdsp::g_tune_mutex);
    for (const auto &kv : blasdsp::g_table) {
        const auto &k = kv.first;
        const auto &v = kv.second;
        fprintf(f, "blasdsp-tune-v1 %d %d %d %d %d %d %d %d %d %.3f\n", std::get<0>(k), std::get<1>(k),
                std::get<2>(k), std::get<3>(k), std::get<4>(k), std::get<5>(k), v.algo, v.tile, v.splitk, v.us);
    }
    return fclose(f) == 0 ? BLASDSP_SUCCESS : BLASDSP_ERR_INVALID;
}

int blasdsp_tune_load(const char *path)
{
    if (!path) return BLASDSP_ERR_INVALID;
    FILE *f = fopen(path, "r");
    if (!f) return BLASDSP_ERR_INVALID;
    char tag[32];
    int d, ta, tb, mb, nb, kb, algo, tile, sk;
    double us;
    int st = BLASDSP_SUCCESS;
    std::map<blasdsp::Key, blasdsp::TuneChoice> loaded;
    for (;;) {
        const int n = fscanf(f, "%31s %d %d %d %d %d %d %d %d %d %lf", tag, &d, &ta, &tb, &mb, &nb, &kb, &algo,
                             &tile, &sk, &us);
        if (n == EOF) break;
        const bool ok = n == 11 && strcmp(tag, "blasdsp-tune-v1") == 0 && (d == 0 || d == 1) &&
                        (ta == 0 || ta == 1) && (tb == 0 || tb == 1) && mb >= 0 && nb >= 0 && kb >= 0 &&
                        ((algo == BLASDSP_ALGO_TILE && (tile == 64 || tile == 128) && sk >= 1 && sk <= 256) ||
                         (algo == BLASDSP_ALGO_SKINNY && sk >= 1 && sk <= 256));
        if (!ok) {
            st = BLASDSP_ERR_INVALID;
            break;
        }
        loaded[blasdsp::Key(d, ta, tb, mb, nb, kb)] = blasdsp::TuneChoice{algo, tile, sk, us};
    }
    fclose(f);
    if (st != BLASDSP_SUCCESS) return st;
    std::lock_guard<std::mutex> lock(blasdsp::g_tune_mutex);
    for (const auto &kv : loaded) blasdsp::g_table[kv.first] = kv.second;
    return BLASDSP_SUCCESS;
}

int blasdsp_tune_clear(void)
{
    std::lock_guard<std::mutex> lock(blasdsp::g_tune_mutex);
    blasdsp::g_table.clear();
    return BLASDSP_SUCCESS;
}

}  // extern "C"
