// grouped.h — device problem list of the grouped persistent GEMM.
#pragma once

#include <vector>

#include "internal.h"

namespace blasdsp {

// One GEMM of a group, as the grouped kernel reads it.  Filled on the host
// from a validated blasdsp_gemm_desc; tiles_n / tile_begin by build_group.
struct GroupProblem {
    const void *A;
    const void *B;
    void *C;
    int64_t M, N, K;
    int64_t lda, ldb, ldc;
    double alpha, beta;
    int64_t tile_begin;   // first tile of this problem in its precision's list
    int64_t tiles_n;
    int32_t bm;           // tile edge, 64 or 128
    uint8_t ta, tb;
};

// Four problem lists, one per grouped kernel, each holding every transpose
// mode: 0: FP32, 128 tile   1: FP32, 64 tile   2: FP64, 64 tile
// 3: FP64, 128 tile.  A process gets four hardware queues, so the four kernel
// nodes of a group's graph all start together.
// Problems of list w are probs[first[w] .. first[w] + count[w]).
constexpr int kGroupLists = 4;
__host__ __device__ inline int group_list(bool f64, int bm)
{
    if (!f64) return bm == 64 ? 1 : 0;
    return bm == 64 ? 2 : 3;
}

// next[w]: the tile counter of list w.  Each workgroup of list w's kernel
// takes its next tile with an atomic increment, so a workgroup that starts
// late (its slot held by another list's kernel) takes what is left instead of
// a fixed share.  The host uploads the header with every counter at zero
// before each replay.
struct GroupHeader {
    int32_t count[kGroupLists];
    int32_t first[kGroupLists];
    int64_t total_tiles[kGroupLists];
    unsigned long long next[kGroupLists];
};

// Persistent grid of list w's kernel (its occupancy times the CU count).
int grouped_grid(int w, int num_cus);
// Threads per workgroup of list w's kernel.
int grouped_block(int w);
// Kernel of list w.
void *grouped_kernel_ptr(int w);
// Sorts each list by decreasing K, assigns tile ranges and writes header
// (counters at zero) and problems; returns the total tile count, or -1 past
// `capacity` problems.
int64_t build_group(const std::vector<GroupProblem> (&lists)[kGroupLists], GroupHeader *hdr,
                    GroupProblem *out, int capacity);

}  // namespace blasdsp
