// serve.hip — continuous-batching GEMM request loop (no counterpart in the
// reference, whose host harness runs one GEMM per launch and waits for it,
// dgemm_release_V1.1/host_code/dgemm_demo.c:148-158).
//
// One host thread drives the loop:
//   arrival   requests become visible when the loop clock passes their
//             (scaled) arrival time;
//   admission FIFO: the next waiting request is admitted when its operands
//             (A, B, C) fit in the HBM budget arena; a request larger than
//             the whole budget is rejected.  Admitted operands stay resident
//             until the request completes;
//   launch    whenever a launch slot is free, every admitted request not yet
//             launched (up to max_batch) goes into one launch on that slot's
//             stream: a grouped fill kernel generates the new operands on the
//             device, the slot's group plan (a hipGraph of the grouped GEMM
//             kernels) runs all of them, a grouped reduction writes partial
//             sums of each C, and they are copied to pinned memory (the
//             host adds a request's partials in a fixed order: its checksum).  With
//             several slots, launches overlap on the device;
//   complete  an event per launch; when it has fired, each request of the
//             launch records its completion time, its memory returns to the
//             arena and newly fitting requests are admitted.
// With grouped = 0 the same loop issues one GEMM launch per request (through
// blasdsp_dgemm / blasdsp_sgemm), the baseline the grouped loop is measured
// against.
#include <math.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <thread>
#include <vector>

#include "api_common.h"

namespace blasdsp {
namespace {

// ---- device helpers ---------------------------------------------------------

struct FillJob {
    void *X;
    int64_t rows, cols;
    uint64_t key;
    int f64;
};

__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ULL;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBULL;
    x ^= x >> 31;
    return x;
}

// U[-1, 1) fill of dense matrices (rows x cols elements in storage order),
// one job per blockIdx.y; same draws as the fill kernel of fill.hip.
__global__ void fill_jobs_kernel(const FillJob *__restrict__ jobs)
{
    const FillJob j = jobs[blockIdx.y];
    const int64_t total = j.rows * j.cols;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t bits = mix64(j.key + ((uint64_t)e + 1) * 0x9E3779B97F4A7C15ULL);
        const double v = 2.0 * ((double)(bits >> 11) * 0x1.0p-53) - 1.0;
        if (j.f64) ((double *)j.X)[e] = v;
        else ((float *)j.X)[e] = (float)v;
    }
}

struct SumJob {
    const void *C;
    int64_t count;
    int f64;
};

constexpr int kSumParts = 64;  // partial sums per request, added on the host in order

// FP64 partial sums of each job's elements: job blockIdx.y, part blockIdx.x,
// fixed summation order (deterministic).
__global__ void checksum_kernel(const SumJob *__restrict__ jobs, double *__restrict__ out)
{
    __shared__ double part[256];
    const SumJob j = jobs[blockIdx.y];
    double s = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < j.count;
         e += (int64_t)gridDim.x * blockDim.x)
        s += j.f64 ? ((const double *)j.C)[e] : (double)((const float *)j.C)[e];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.y * kSumParts + blockIdx.x] = part[0];
}

uint64_t key_of(uint64_t seed)
{
    uint64_t k = seed * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL;
    k ^= k >> 30;
    k *= 0xBF58476D1CE4E5B9ULL;
    k ^= k >> 27;
    k *= 0x94D049BB133111EBULL;
    k ^= k >> 31;
    return k;
}

// ---- HBM arena: first fit over a free map, 256-byte granules -----------------

class Arena {
  public:
    explicit Arena(int64_t bytes) : size_(bytes) { free_[0] = bytes; }
    int64_t alloc(int64_t bytes)
    {
        bytes = (bytes + 255) / 256 * 256;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                const int64_t off = it->first, rest = it->second - bytes;
                free_.erase(it);
                if (rest > 0) free_[off + bytes] = rest;
                used_ += bytes;
                peak_ = std::max(peak_, used_);
                return off;
            }
        }
        return -1;
    }
    void release(int64_t off, int64_t bytes)
    {
        bytes = (bytes + 255) / 256 * 256;
        used_ -= bytes;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + bytes == next->first) {
            bytes += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += bytes;
                return;
            }
        }
        free_[off] = bytes;
    }
    int64_t size() const { return size_; }
    int64_t peak() const { return peak_; }

  private:
    int64_t size_, used_ = 0, peak_ = 0;
    std::map<int64_t, int64_t> free_;
};

// ---- loop ------------------------------------------------------------------

struct Live {
    int idx;
    int64_t off, bytes;
    void *A, *B, *C;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    blasdsp_group_plan *plan = nullptr;
    FillJob *h_fill = nullptr, *d_fill = nullptr;
    SumJob *h_sum = nullptr, *d_sum = nullptr;
    double *h_out = nullptr, *d_out = nullptr;
    std::vector<Live> reqs;
    int batch_id = -1;
    bool busy = false;
};

double now_s()
{
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int64_t elem_bytes(const blasdsp_request &r) { return r.dtype == BLASDSP_FLOAT64 ? 8 : 4; }

int64_t req_bytes(const blasdsp_request &r)
{
    const int64_t s = elem_bytes(r);
    auto up = [](int64_t b) { return (b + 255) / 256 * 256; };
    return up(s * r.M * r.K) + up(s * r.K * r.N) + up(s * r.M * r.N);
}

bool valid(const blasdsp_request &r)
{
    bool t;
    return (r.dtype == BLASDSP_FLOAT32 || r.dtype == BLASDSP_FLOAT64) && parse_trans(r.transA, t) &&
           parse_trans(r.transB, t) && r.M > 0 && r.N > 0 && r.K > 0;
}

blasdsp_gemm_desc desc_of(const blasdsp_request &r, const Live &l)
{
    blasdsp_gemm_desc d;
    bool ta = false, tb = false;
    parse_trans(r.transA, ta);
    parse_trans(r.transB, tb);
    d.dtype = r.dtype;
    d.transA = ta ? 'T' : 'N';
    d.transB = tb ? 'T' : 'N';
    d.M = r.M; d.N = r.N; d.K = r.K;
    d.alpha = r.alpha; d.beta = r.beta;
    d.A = l.A; d.lda = ta ? r.M : r.K;   // dense operands
    d.B = l.B; d.ldb = tb ? r.K : r.N;
    d.C = l.C; d.ldc = r.N;
    return d;
}

void free_slot(Slot &s)
{
    if (s.plan) blasdsp_group_plan_destroy(s.plan);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.stream) {
        // the per-request path may have given this stream a split-K workspace
        (void)workspace_release(s.stream);
        (void)hipStreamDestroy(s.stream);
    }
    if (s.h_fill) (void)hipHostFree(s.h_fill);
    if (s.h_sum) (void)hipHostFree(s.h_sum);
    if (s.h_out) (void)hipHostFree(s.h_out);
    if (s.d_fill) (void)hipFree(s.d_fill);
    if (s.d_sum) (void)hipFree(s.d_sum);
    if (s.d_out) (void)hipFree(s.d_out);
}

hipError_t init_slot(Slot &s, int max_batch)
{
    hipError_t e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipHostMalloc((void **)&s.h_fill, sizeof(FillJob) * 3 * max_batch, 0);
    if (e == hipSuccess) e = hipHostMalloc((void **)&s.h_sum, sizeof(SumJob) * max_batch, 0);
    if (e == hipSuccess) e = hipHostMalloc((void **)&s.h_out, sizeof(double) * kSumParts * max_batch, 0);
    if (e == hipSuccess) e = hipMalloc((void **)&s.d_fill, sizeof(FillJob) * 3 * max_batch);
    if (e == hipSuccess) e = hipMalloc((void **)&s.d_sum, sizeof(SumJob) * max_batch);
    if (e == hipSuccess) e = hipMalloc((void **)&s.d_out, sizeof(double) * kSumParts * max_batch);
    if (e != hipSuccess) return e;
    if (blasdsp_group_plan_create(max_batch, &s.plan) != BLASDSP_SUCCESS) return hipErrorOutOfMemory;
    return hipSuccess;
}

// Enqueue one launch of s.reqs on s.stream.
int launch(Slot &s, const blasdsp_request *reqs, bool grouped)
{
    const int n = (int)s.reqs.size();
    for (int i = 0; i < n; ++i) {
        const Live &l = s.reqs[i];
        const blasdsp_request &r = reqs[l.idx];
        const int f64 = r.dtype == BLASDSP_FLOAT64;
        // Stored A, B, C (dense) take the generator streams of seeds 3 seed,
        // 3 seed + 1 and 3 seed + 2: element e of the stored matrix is draw e,
        // as blasdsp_fill_* would write it with ld = stored columns.
        s.h_fill[3 * i + 0] = FillJob{l.A, r.M, r.K, key_of(3 * r.seed + 0), f64};
        s.h_fill[3 * i + 1] = FillJob{l.B, r.K, r.N, key_of(3 * r.seed + 1), f64};
        s.h_fill[3 * i + 2] = FillJob{l.C, r.M, r.N, key_of(3 * r.seed + 2), f64};
        s.h_sum[i] = SumJob{l.C, r.M * r.N, f64};
    }
    hipError_t e = hipMemcpyAsync(s.d_fill, s.h_fill, sizeof(FillJob) * 3 * n, hipMemcpyHostToDevice, s.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s.d_sum, s.h_sum, sizeof(SumJob) * n, hipMemcpyHostToDevice, s.stream);
    if (e != hipSuccess) return hip_status(e);
    hipLaunchKernelGGL(fill_jobs_kernel, dim3(256, 3 * n), dim3(256), 0, s.stream, s.d_fill);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_status(e);
    if (grouped) {
        std::vector<blasdsp_gemm_desc> d(n);
        for (int i = 0; i < n; ++i) d[i] = desc_of(reqs[s.reqs[i].idx], s.reqs[i]);
        const int st = blasdsp_group_plan_run(s.plan, d.data(), n, s.stream);
        if (st != BLASDSP_SUCCESS) return st;
    } else {
        for (int i = 0; i < n; ++i) {
            const blasdsp_gemm_desc d = desc_of(reqs[s.reqs[i].idx], s.reqs[i]);
            const int st = d.dtype == BLASDSP_FLOAT64
                               ? blasdsp_dgemm(d.transA, d.transB, d.M, d.N, d.K, d.alpha, (const double *)d.A,
                                               d.lda, (const double *)d.B, d.ldb, d.beta, (double *)d.C, d.ldc,
                                               s.stream)
                               : blasdsp_sgemm(d.transA, d.transB, d.M, d.N, d.K, (float)d.alpha,
                                               (const float *)d.A, d.lda, (const float *)d.B, d.ldb,
                                               (float)d.beta, (float *)d.C, d.ldc, s.stream);
            if (st != BLASDSP_SUCCESS) return st;
        }
    }
    hipLaunchKernelGGL(checksum_kernel, dim3(kSumParts, n), dim3(256), 0, s.stream, s.d_sum, s.d_out);
    e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(s.h_out, s.d_out, sizeof(double) * kSumParts * n, hipMemcpyDeviceToHost, s.stream);
    if (e == hipSuccess) e = hipEventRecord(s.done, s.stream);
    return hip_status(e);
}

}  // namespace
}  // namespace blasdsp

extern "C" int blasdsp_serve(const blasdsp_request *reqs, int n, const blasdsp_serve_config *cfg,
                             blasdsp_request_result *results, blasdsp_serve_stats *stats, void **outputs)
{
    using namespace blasdsp;
    if (n < 0 || (n > 0 && (!reqs || !results)) || !cfg) return BLASDSP_ERR_INVALID;
    // max_batch bounds the grouped fill launch (3 jobs per request in grid.y <= 65535)
    if (cfg->memory_budget <= 0 || cfg->max_batch <= 0 || cfg->max_batch > 16384 || cfg->max_inflight <= 0 ||
        cfg->max_inflight > 16 || cfg->time_scale < 0)
        return BLASDSP_ERR_INVALID;
    const bool grouped = cfg->grouped != 0;
    const int max_batch = grouped ? cfg->max_batch : 1;

    void *base = nullptr;
    hipError_t e = hipMalloc(&base, (size_t)cfg->memory_budget);
    if (e != hipSuccess) return hip_status(e);
    Arena arena(cfg->memory_budget);
    std::vector<Slot> slots(cfg->max_inflight);
    for (Slot &s : slots) {
        e = init_slot(s, max_batch);
        if (e != hipSuccess) {
            for (Slot &t : slots) free_slot(t);
            (void)hipFree(base);
            return hip_status(e);
        }
    }

    // arrival order
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    for (int i = 0; i < n; ++i) {
        results[i] = blasdsp_request_result{};
        results[i].status = 1;  // pending
    }

    std::deque<int> waiting;     // arrived, not admitted
    std::deque<Live> admitted;   // resident, not launched
    int next = 0, finished = 0, launches = 0;
    int64_t launched_reqs = 0;
    int status = BLASDSP_SUCCESS;
    double flops = 0.0;
    const double t0 = now_s();

    while (finished < n && status == BLASDSP_SUCCESS) {
        bool progress = false;
        const double t = now_s() - t0;
        // arrivals
        while (next < n && reqs[order[next]].arrival_s * cfg->time_scale <= t) {
            waiting.push_back(order[next++]);
            progress = true;
        }
        // admission (FIFO)
        while (!waiting.empty()) {
            const int i = waiting.front();
            const blasdsp_request &r = reqs[i];
            if (!valid(r)) {
                results[i].status = BLASDSP_ERR_INVALID;
                waiting.pop_front();
                ++finished;
                progress = true;
                continue;
            }
            const int64_t bytes = req_bytes(r);
            if (bytes > arena.size()) {
                results[i].status = BLASDSP_ERR_NOMEM;
                waiting.pop_front();
                ++finished;
                progress = true;
                continue;
            }
            const int64_t off = arena.alloc(bytes);
            if (off < 0) break;
            char *p = (char *)base + off;
            const int64_t s = elem_bytes(r);
            auto up = [](int64_t b) { return (b + 255) / 256 * 256; };
            Live l{i, off, bytes, p, p + up(s * r.M * r.K), p + up(s * r.M * r.K) + up(s * r.K * r.N)};
            results[i].start_s = now_s() - t0;
            admitted.push_back(l);
            waiting.pop_front();
            progress = true;
        }
        // completions
        for (Slot &s : slots) {
            if (!s.busy) continue;
            hipError_t q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
                status = hip_status(q);
                break;
            }
            const double td = now_s() - t0;
            for (size_t k = 0; k < s.reqs.size(); ++k) {
                const Live &l = s.reqs[k];
                const blasdsp_request &r = reqs[l.idx];
                blasdsp_request_result &res = results[l.idx];
                res.done_s = td;
                res.latency_s = td - r.arrival_s * cfg->time_scale;
                double sum = 0.0;
                for (int q = 0; q < kSumParts; ++q) sum += s.h_out[k * kSumParts + q];
                res.checksum = sum;
                res.batch = s.batch_id;
                res.status = BLASDSP_SUCCESS;
                if (outputs && outputs[l.idx]) {
                    const hipError_t ce = hipMemcpy(outputs[l.idx], l.C, (size_t)(elem_bytes(r) * r.M * r.N),
                                                    hipMemcpyDeviceToHost);
                    if (ce != hipSuccess) res.status = hip_status(ce);
                }
                flops += 2.0 * (double)r.M * (double)r.N * (double)r.K;
                arena.release(l.off, l.bytes);
                ++finished;
            }
            s.reqs.clear();
            s.busy = false;
            progress = true;
        }
        if (status != BLASDSP_SUCCESS) break;
        // launches
        for (Slot &s : slots) {
            if (s.busy || admitted.empty()) continue;
            while (!admitted.empty() && (int)s.reqs.size() < max_batch) {
                s.reqs.push_back(admitted.front());
                admitted.pop_front();
            }
            s.batch_id = launches++;
            launched_reqs += (int64_t)s.reqs.size();
            const int st = launch(s, reqs, grouped);
            if (st != BLASDSP_SUCCESS) {
                status = st;
                break;
            }
            s.busy = true;
            progress = true;
        }
        if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(5));
    }
    const double wall = now_s() - t0;
    for (Slot &s : slots) {
        if (s.busy) (void)hipEventSynchronize(s.done);
        free_slot(s);
    }
    (void)hipFree(base);

    if (stats) {
        std::vector<double> lat;
        int64_t rejected = 0;
        for (int i = 0; i < n; ++i) {
            if (results[i].status == BLASDSP_SUCCESS) lat.push_back(results[i].latency_s);
            else ++rejected;
        }
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double q) {
            if (lat.empty()) return 0.0;
            size_t k = (size_t)ceil(q * (double)lat.size());
            if (k < 1) k = 1;
            return lat[std::min(k, lat.size()) - 1];
        };
        double sum = 0.0;
        for (double v : lat) sum += v;
        memset(stats, 0, sizeof(*stats));
        stats->completed = (int64_t)lat.size();
        stats->rejected = rejected;
        stats->wall_s = wall;
        stats->gemms_per_s = wall > 0 ? (double)lat.size() / wall : 0.0;
        stats->tflops = wall > 0 ? flops / wall / 1e12 : 0.0;
        stats->p50_s = pct(0.50);
        stats->p99_s = pct(0.99);
        stats->mean_s = lat.empty() ? 0.0 : sum / (double)lat.size();
        stats->max_s = lat.empty() ? 0.0 : lat.back();
        stats->launches = launches;
        stats->mean_batch = launches > 0 ? (double)launched_reqs / (double)launches : 0.0;
        stats->peak_resident_bytes = arena.peak();
    }
    return status;
}
