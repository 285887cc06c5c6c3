// gemm_grouped.hip — many GEMMs of uneven shapes and transpose modes in one
// replay of four persistent kernels.
//
// No counterpart in the reference (which runs one GEMM per hthread_group_exec,
// dgemm_release_V1.1/host_code/dgemm_demo.c:151).  The host sorts the group's
// problems into four lists by (precision, tile edge): FP32 and FP64, 128 x 128
// or 64 x 64 tiles, each list holding every transpose mode.  It flattens each
// list into C tiles: problem p owns tiles [tile_begin(p), tile_begin(p) +
// tiles_m * tiles_n), problems sorted by K (the cost of one of their tiles) in
// decreasing order.  Each list has its own kernel with the register budget and
// LDS of its tile size; a fixed grid of workgroups (occupancy times CU count)
// takes tiles from the list's atomic counter one at a time, so the expensive
// tiles go first and a workgroup that starts late (its CU slot held by
// another list's kernel) takes whatever is left.  A workgroup finds its
// problem by binary search over tile_begin and runs the same tile body as the
// per-call kernel (tile_body.h), so a problem's C is bit-identical alone and
// in a group.  The four kernels are independent graph nodes; a process has
// four hardware queues, so they start together and share the device.
//
// The problem list lives in device memory and the grids do not depend on it,
// so an instantiated hipGraph (built from explicit kernel nodes, no stream
// capture and no memset nodes) replays any group that fills its lists: the
// host uploads the list, with the tile counters at zero, and launches the
// graph of the lists that have problems.
//
// On 512 requests of a mixed workload (M, N, K log-uniform in [16, 4096], 30 %
// with one side 8-32, 70 % FP32, every mode, unaligned leading dimensions) on
// MI355X: seven kernels (FP64 128 tiles one per mode) with a fixed stride
// over the tiles took 5.8 ms, the three FP64 128 kernels starting only when
// one of the first four had finished; four kernels with the tile counter 5.4
// ms; loads at any alignment in the tile body 4.5 ms.  Tiles handed out in
// list order go to whichever workgroup asks, so the tiles one XCD runs
// together share few A and B panels (L2 hit rate 0.40 on one 4095 x 4096 x
// 4096 FP64 problem, 0.74 for the per-call kernel's XCD-aware order).  Per-XCD
// tile queues (blocks of 8 x 8 tiles dealt to the XCDs, workgroups stealing
// from the other queues once their own is empty) raised it to 0.73 and ran no
// faster: 4.59 against 4.50 ms on the mixed workload.
#include <stdlib.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "grouped.h"
#include "tile_body.h"

namespace blasdsp {
namespace {

using tile::Cfg;

// Waves per workgroup: 128 tiles take the 8-wave body (32 x 64 per wave) of
// the per-call FP64 kernel, 64 tiles 4 waves (32 x 32 each).  On the mixed
// workload below, with one kernel per mode, 8 waves for FP64 ran 3 % faster
// than 4, and 8 waves for both precisions 7 % faster than 4 for both.  With
// all modes in one kernel: 4 waves for FP64 ran uniform 2048^3 groups 6-30 %
// faster but the mixed workload's FP64 128 list 4 % slower; 4 waves for FP32
// (the per-call FP32 kernel's 64 x 64 per wave) ran one 4095 x 4096 x 4096 NN
// problem 16 % faster but the mixed workload's FP32 128 list 9 % slower.
template <typename T, int BM>
constexpr int group_nw()
{
    return BM == 128 ? 8 : 4;
}

template <typename T, int BM, bool TA, bool TB>
__device__ __forceinline__ void run_tile(const GroupProblem &P, int64_t local, T *lds, int tid)
{
    Gemm<T> g;
    g.ta = TA; g.tb = TB;
    g.M = P.M; g.N = P.N; g.K = P.K;
    g.alpha = (T)P.alpha; g.beta = (T)P.beta;
    g.A = (const T *)P.A; g.lda = P.lda;
    g.B = (const T *)P.B; g.ldb = P.ldb;
    g.C = (T *)P.C; g.ldc = P.ldc;
    const int64_t tm = local / P.tiles_n, tn = local % P.tiles_n;
    tile::tile_body<T, BM, TA, TB, group_nw<T, BM>()>(g, tm * BM, tn * BM, lds, 0, -1, nullptr, tid);
}

template <typename T, int BM>
__global__ void __launch_bounds__((64 * group_nw<T, BM>()))
__attribute__((amdgpu_waves_per_eu(Cfg<T, BM, group_nw<T, BM>()>::WPE, Cfg<T, BM, group_nw<T, BM>()>::WPE)))
grouped_kernel(GroupHeader *__restrict__ hdr, const GroupProblem *__restrict__ probs)
{
    // The tile number is passed to the waves through the first 8 bytes of the
    // stage buffers (a separate LDS variable would cost the 64-tile kernels a
    // workgroup per CU), so it is read by every wave before the body writes.
    __shared__ __attribute__((aligned(16))) T lds[Cfg<T, BM, group_nw<T, BM>()>::LDS_ELEMS];
    unsigned long long *const s_tile = reinterpret_cast<unsigned long long *>(lds);
    const int w = group_list(sizeof(T) == 8, BM);
    const int count = hdr->count[w];
    const int64_t total = hdr->total_tiles[w];
    probs += hdr->first[w];
    for (;;) {
        if (threadIdx.x == 0) *s_tile = atomicAdd(&hdr->next[w], 1ull);
        __syncthreads();
        // Read through readfirstlane: the value is the same in every lane, and
        // a scalar tile number keeps the problem lookup and the problem's
        // fields in scalar registers.
        const unsigned long long tv = *s_tile;
        const int64_t t = (int64_t)(((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(tv >> 32)) << 32) |
                                    (unsigned)__builtin_amdgcn_readfirstlane((unsigned)tv));
        __syncthreads();
        if (t >= total) break;  // the same value in every wave: all leave together
        int lo = 0, hi = count - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (probs[mid].tile_begin <= t) lo = mid;
            else hi = mid - 1;
        }
        const GroupProblem &P = probs[lo];
        const int64_t local = t - P.tile_begin;
        // FP64 128 tiles: an opaque copy of the thread index, so the
        // per-thread offsets derived from it are computed inside each tile
        // instead of being kept live across the loop.  Kept live they spilled
        // (80 bytes per lane), and the reloads inside the K loop waited on the
        // tile's global loads: 2.10 -> 1.91 ms for the mixed workload's FP64
        // 128 list, 2.85 -> 2.51 ms for eight 2048^3 FP64 GEMMs.  The other
        // three kernels do not spill and ran 5-18 % slower with it.
        int tid = threadIdx.x;
        if constexpr (sizeof(T) == 8 && BM == 128) asm volatile("" : "+v"(tid));
        switch ((P.ta << 1) | P.tb) {
        case 0: run_tile<T, BM, false, false>(P, local, lds, tid); break;
        case 1: run_tile<T, BM, false, true>(P, local, lds, tid); break;
        case 2: run_tile<T, BM, true, false>(P, local, lds, tid); break;
        default: run_tile<T, BM, true, true>(P, local, lds, tid); break;
        }
        __syncthreads();  // the next tile number and stages reuse the LDS
    }
}

}  // namespace

int grouped_grid(int w, int num_cus)
{
    const int cus = num_cus > 0 ? num_cus : 256;
    int wpe, nw;
    switch (w) {
    case 0: wpe = Cfg<float, 128, group_nw<float, 128>()>::WPE; nw = group_nw<float, 128>(); break;
    case 1: wpe = Cfg<float, 64, group_nw<float, 64>()>::WPE; nw = group_nw<float, 64>(); break;
    case 2: wpe = Cfg<double, 64, group_nw<double, 64>()>::WPE; nw = group_nw<double, 64>(); break;
    default: wpe = Cfg<double, 128, group_nw<double, 128>()>::WPE; nw = group_nw<double, 128>(); break;
    }
    return cus * wpe * 4 / nw;  // resident workgroups: waves per SIMD x 4 SIMDs / waves per workgroup
}

int grouped_block(int w)
{
    switch (w) {
    case 0: return 64 * group_nw<float, 128>();
    case 1: return 64 * group_nw<float, 64>();
    case 2: return 64 * group_nw<double, 64>();
    default: return 64 * group_nw<double, 128>();
    }
}

void *grouped_kernel_ptr(int w)
{
    switch (w) {
    case 0: return reinterpret_cast<void *>(&grouped_kernel<float, 128>);
    case 1: return reinterpret_cast<void *>(&grouped_kernel<float, 64>);
    case 2: return reinterpret_cast<void *>(&grouped_kernel<double, 64>);
    default: return reinterpret_cast<void *>(&grouped_kernel<double, 128>);
    }
}

int64_t build_group(const std::vector<GroupProblem> (&lists)[kGroupLists], GroupHeader *hdr,
                    GroupProblem *out, int capacity)
{
    int pos = 0;
    int64_t tiles_all = 0;
    for (int w = 0; w < kGroupLists; ++w) {
        std::vector<GroupProblem> v = lists[w];
        std::stable_sort(v.begin(), v.end(),
                         [](const GroupProblem &a, const GroupProblem &b) { return a.K > b.K; });
        hdr->first[w] = pos;
        int64_t tiles = 0;
        for (GroupProblem p : v) {
            if (pos >= capacity) return -1;
            const int64_t bm = p.bm;
            const int64_t tm = (p.M + bm - 1) / bm;
            p.tiles_n = (p.N + bm - 1) / bm;
            p.tile_begin = tiles;
            tiles += tm * p.tiles_n;
            out[pos++] = p;
        }
        hdr->count[w] = (int)v.size();
        hdr->total_tiles[w] = tiles;
        hdr->next[w] = 0;
        tiles_all += tiles;
    }
    return tiles_all;
}

}  // namespace blasdsp
