// group_api.hip — C ABI of the grouped persistent GEMM and its hipGraph plan.
//
// A plan owns the device problem list, a pinned staging copy of it and one
// instantiated graph per non-empty set of problem lists: the graph of set S
// holds one kernel node per list in S (no edges between them), each at its
// full persistent grid.  The instantiated graphs are never changed after
// instantiation: running a plan validates the descriptors, builds the sorted
// tile list (tile counters at zero) in the staging buffer, uploads it with one
// async copy on the caller's stream and replays, on that stream, the graph of
// the lists that received problems.  A group of one precision and tile size
// thus runs one kernel: with the idle lists' kernels in the same replay (their
// workgroups take no tile and exit) one 4095 x 4096 x 4096 FP64 problem took
// 2.17-2.65 ms a replay, alone 2.17-2.18 ms.  The staging buffer is reused
// only after the previous upload has completed (an event recorded after the
// copy).
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>
#include <new>
#include <utility>
#include <vector>

#include "api_common.h"
#include "grouped.h"
#include "tile_body.h"
#include "tune.h"

struct blasdsp_group_plan {
    int device = 0;
    int capacity = 0;
    size_t list_bytes = 0;        // header slot + capacity problems
    char *d_list = nullptr;       // device: [GroupHeader | pad][GroupProblem x capacity]
    char *h_list = nullptr;       // pinned staging, same layout
    hipEvent_t uploaded = nullptr;  // the staging copy has been read
    hipEvent_t done = nullptr;      // the last graph replay has finished
    bool pending_upload = false;
    bool has_run = false;           // `done` has been recorded
    // graph[S], exec[S]: the kernels of the lists in bit set S (S = 1 .. 15)
    hipGraph_t graph[1 << blasdsp::kGroupLists] = {};
    hipGraphExec_t exec[1 << blasdsp::kGroupLists] = {};
    int full_grid[blasdsp::kGroupLists] = {};
};

namespace blasdsp {
namespace {

constexpr size_t kHeaderSlot = 256;
static_assert(sizeof(GroupHeader) <= kHeaderSlot, "header fits its slot");

GroupHeader *hdr_of(char *base) { return reinterpret_cast<GroupHeader *>(base); }
GroupProblem *probs_of(char *base) { return reinterpret_cast<GroupProblem *>(base + kHeaderSlot); }

// Kernel node parameters of list w.  The argument values are copied when the
// node is added or updated, so the locals behind kernelParams may go away.
struct NodeArgs {
    GroupHeader *hdr;
    const GroupProblem *probs;
    void *ptrs[2];
};

hipKernelNodeParams node_params(blasdsp_group_plan *p, int w, int grid)
{
    static thread_local NodeArgs a;
    a.hdr = hdr_of(p->d_list);
    a.probs = probs_of(p->d_list);
    a.ptrs[0] = (void *)&a.hdr;
    a.ptrs[1] = (void *)&a.probs;
    hipKernelNodeParams kp{};
    kp.func = grouped_kernel_ptr(w);
    kp.gridDim = dim3(grid);
    kp.blockDim = dim3(grouped_block(w));
    kp.sharedMemBytes = 0;
    kp.kernelParams = a.ptrs;
    kp.extra = nullptr;
    return kp;
}

void destroy_plan(blasdsp_group_plan *p)
{
    if (!p) return;
    for (int m = 0; m < (1 << kGroupLists); ++m) {
        if (p->exec[m]) (void)hipGraphExecDestroy(p->exec[m]);
        if (p->graph[m]) (void)hipGraphDestroy(p->graph[m]);
    }
    if (p->uploaded) (void)hipEventDestroy(p->uploaded);
    if (p->done) (void)hipEventDestroy(p->done);
    if (p->h_list) (void)hipHostFree(p->h_list);
    if (p->d_list) (void)hipFree(p->d_list);
    delete p;
}

int create_plan(int capacity, blasdsp_group_plan **out)
{
    if (capacity <= 0 || !out) return BLASDSP_ERR_INVALID;
    blasdsp_group_plan *p = new (std::nothrow) blasdsp_group_plan();
    if (!p) return BLASDSP_ERR_NOMEM;
    hipError_t e = hipGetDevice(&p->device);
    p->capacity = capacity;
    p->list_bytes = kHeaderSlot + (size_t)capacity * sizeof(GroupProblem);
    if (e == hipSuccess) e = hipMalloc(&p->d_list, p->list_bytes);
    if (e == hipSuccess) e = hipHostMalloc(&p->h_list, p->list_bytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->uploaded, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->done, hipEventDisableTiming);
    if (e != hipSuccess) {
        destroy_plan(p);
        return hip_status(e);
    }
    memset(p->h_list, 0, p->list_bytes);
    const int cus = num_cus();
    for (int w = 0; w < kGroupLists; ++w) p->full_grid[w] = grouped_grid(w, cus);
    for (int m = 1; m < (1 << kGroupLists) && e == hipSuccess; ++m) {
        e = hipGraphCreate(&p->graph[m], 0);
        for (int w = 0; w < kGroupLists && e == hipSuccess; ++w) {
            if (!(m >> w & 1)) continue;
            hipGraphNode_t node;
            hipKernelNodeParams kp = node_params(p, w, p->full_grid[w]);
            e = hipGraphAddKernelNode(&node, p->graph[m], nullptr, 0, &kp);
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&p->exec[m], p->graph[m], nullptr, nullptr, 0);
    }
    if (e != hipSuccess) {
        destroy_plan(p);
        return hip_status(e);
    }
    *out = p;
    return BLASDSP_SUCCESS;
}

// Tile edge of one problem of a group.  A group has parallelism across its
// problems, so the choice is the cheaper cover of the problem's C: tiles x
// relative tile cost, a 128 tile costing kTile128Cost 64 tiles (its four
// times the work runs on a larger, more efficient per-wave tile).  On 512
// requests of a mixed workload (M, N, K log-uniform in [16, 4096]) costs 2.8
// to 3.2 ran 1.5 % faster than 128 tiles only when M and N exceed 96; 3.6
// was 2.5 % slower, 4.0 35 % slower.
constexpr double kTile128Cost = 3.0;
int group_tile(int64_t M, int64_t N)
{
    const double c64 = (double)((M + 63) / 64) * (double)((N + 63) / 64);
    const double c128 = (double)((M + 127) / 128) * (double)((N + 127) / 128) * kTile128Cost;
    return c128 <= c64 ? 128 : 64;
}

// direct: the grouped kernel's buffer loads do not take the problem (leading
// dimensions past their 32-bit offsets, or a stored operand narrower than one
// 16-byte vector: tile::buffer_ok); it runs through the per-call tile kernel
// on the plan's stream instead (same C bits).
template <typename T>
int to_problem(const blasdsp_gemm_desc &d, GroupProblem &p, bool &skip, bool &direct)
{
    Gemm<T> g;
    int st = check<T>(g, d.transA, d.transB, d.M, d.N, d.K, (T)d.alpha, (const T *)d.A, d.lda,
                      (const T *)d.B, d.ldb, (T)d.beta, (T *)d.C, d.ldc);
    if (st != BLASDSP_SUCCESS) return st;
    skip = (d.M == 0 || d.N == 0);
    p = GroupProblem{};
    p.A = d.A; p.B = d.B; p.C = d.C;
    p.M = d.M; p.N = d.N;
    p.K = ((T)d.alpha == T(0)) ? 0 : d.K;  // alpha == 0: C = beta C, A and B not read
    p.lda = d.lda; p.ldb = d.ldb; p.ldc = d.ldc;
    p.alpha = d.alpha; p.beta = d.beta;
    p.ta = g.ta; p.tb = g.tb;
    TuneChoice c;
    if (tune_lookup(sizeof(T) == 8, g.ta, g.tb, d.M, d.N, d.K, &c) && c.algo == BLASDSP_ALGO_TILE)
        p.bm = c.tile;
    else
        p.bm = group_tile(d.M, d.N);
    direct = !tile::buffer_ok<T, 128>(g);
    return BLASDSP_SUCCESS;
}

template <typename T>
hipError_t run_direct(const GroupProblem &p, hipStream_t s)
{
    Gemm<T> g;
    g.ta = p.ta; g.tb = p.tb;
    g.M = p.M; g.N = p.N; g.K = p.K;
    g.alpha = (T)p.alpha; g.beta = (T)p.beta;
    g.A = (const T *)p.A; g.lda = p.lda;
    g.B = (const T *)p.B; g.ldb = p.ldb;
    g.C = (T *)p.C; g.ldc = p.ldc;
    return launch_tile<T>(g, p.bm, 1, nullptr, s);
}

int run_plan(blasdsp_group_plan *p, const blasdsp_gemm_desc *descs, int count, hipStream_t s)
{
    if (!p || count < 0 || (count > 0 && !descs)) return BLASDSP_ERR_INVALID;
    if (count > p->capacity) return BLASDSP_ERR_INVALID;
    std::vector<GroupProblem> lists[kGroupLists];
    std::vector<std::pair<bool, GroupProblem>> direct_probs;
    int used = 0;
    for (int i = 0; i < count; ++i) {
        GroupProblem gp;
        bool skip = false, direct = false;
        int st;
        const bool f64 = descs[i].dtype == BLASDSP_FLOAT64;
        if (descs[i].dtype == BLASDSP_FLOAT32) st = to_problem<float>(descs[i], gp, skip, direct);
        else if (f64) st = to_problem<double>(descs[i], gp, skip, direct);
        else st = BLASDSP_ERR_INVALID;
        if (st != BLASDSP_SUCCESS) return st;
        if (skip) continue;
        if (direct) {
            direct_probs.emplace_back(f64, gp);
        } else {
            lists[group_list(f64, gp.bm)].push_back(gp);
            ++used;
        }
    }
    // The staging buffer may still be read by the previous upload.
    if (p->pending_upload) {
        hipError_t e = hipEventSynchronize(p->uploaded);
        if (e != hipSuccess) return hip_status(e);
        p->pending_upload = false;
    }
    GroupHeader *h = hdr_of(p->h_list);
    if (build_group(lists, h, probs_of(p->h_list), p->capacity) < 0) return BLASDSP_ERR_INVALID;
    const size_t bytes = kHeaderSlot + (size_t)used * sizeof(GroupProblem);
    // The previous replay may still be reading the device list and counters:
    // order the copy after it (on the same stream this wait is already met).
    hipError_t e = p->has_run ? hipStreamWaitEvent(s, p->done, 0) : hipSuccess;
    if (e == hipSuccess) e = hipMemcpyAsync(p->d_list, p->h_list, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipEventRecord(p->uploaded, s);
    if (e != hipSuccess) return hip_status(e);
    p->pending_upload = true;
    int mask = 0;
    for (int w = 0; w < kGroupLists; ++w)
        if (!lists[w].empty()) mask |= 1 << w;
    if (mask) e = hipGraphLaunch(p->exec[mask], s);
    for (size_t i = 0; i < direct_probs.size() && e == hipSuccess; ++i)
        e = direct_probs[i].first ? run_direct<double>(direct_probs[i].second, s)
                                  : run_direct<float>(direct_probs[i].second, s);
    if (e == hipSuccess) e = hipEventRecord(p->done, s);
    if (e == hipSuccess) p->has_run = true;
    return hip_status(e);
}

// Plans behind blasdsp_gemm_grouped: one per (device, stream), grown on demand.
std::mutex g_plan_mutex;
std::map<std::pair<int, hipStream_t>, blasdsp_group_plan *> g_plans;

}  // namespace

hipError_t grouped_plan_release(hipStream_t s)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(g_plan_mutex);
    auto it = g_plans.find(std::make_pair(dev, s));
    if (it == g_plans.end()) return hipSuccess;
    if (it->second) {
        e = hipEventSynchronize(it->second->done);  // never recorded: returns at once
        destroy_plan(it->second);
    }
    g_plans.erase(it);
    return e;
}

}  // namespace blasdsp

extern "C" {

int blasdsp_group_plan_create(int max_problems, blasdsp_group_plan **plan)
{
    return blasdsp::create_plan(max_problems, plan);
}

int blasdsp_group_plan_run(blasdsp_group_plan *plan, const blasdsp_gemm_desc *descs, int count,
                           void *stream)
{
    return blasdsp::run_plan(plan, descs, count, (hipStream_t)stream);
}

int blasdsp_group_plan_destroy(blasdsp_group_plan *plan)
{
    if (!plan) return BLASDSP_ERR_INVALID;
    (void)hipEventSynchronize(plan->done);  // never recorded: returns at once
    blasdsp::destroy_plan(plan);
    return BLASDSP_SUCCESS;
}

int blasdsp_gemm_grouped(const blasdsp_gemm_desc *descs, int count, void *stream)
{
    if (count < 0 || (count > 0 && !descs)) return BLASDSP_ERR_INVALID;
    if (count == 0) return BLASDSP_SUCCESS;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return blasdsp::hip_status(e);
    hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lock(blasdsp::g_plan_mutex);
    blasdsp_group_plan *&p = blasdsp::g_plans[std::make_pair(dev, s)];
    if (p && p->capacity < count) {
        // the previous graph may still be running on s
        e = hipEventSynchronize(p->done);
        if (e != hipSuccess) return blasdsp::hip_status(e);
        blasdsp::destroy_plan(p);
        p = nullptr;
    }
    if (!p) {
        int cap = 256;
        while (cap < count) cap *= 2;
        int st = blasdsp::create_plan(cap, &p);
        if (st != BLASDSP_SUCCESS) {
            p = nullptr;
            return st;
        }
    }
    return blasdsp::run_plan(p, descs, count, s);
}

}  // extern "C"
