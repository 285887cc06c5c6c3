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
    uint8_t ta, tb, a_vec, b_vec;
};

// Seven problem lists, one per grouped kernel:
//   0: FP32, 128 tile    1: FP32, 64 tile    2: FP64, 64 tile   (all modes)
//   3 + mode: FP64, 128 tile, one list per transpose mode (ta * 2 + tb)
// The FP64 128 tile body needs ~240 VGPRs; folding its four modes into one
// kernel pushes it past 256 and spills, so that case gets a kernel per mode.
// Problems of list w are probs[first[w] .. first[w] + count[w]).
constexpr int kGroupLists = 7;
__host__ __device__ inline int group_list(bool f64, int bm, int mode)
{
    if (!f64) return bm == 64 ? 1 : 0;
    return bm == 64 ? 2 : 3 + mode;
}

struct GroupHeader {
    int32_t count[kGroupLists];
    int32_t first[kGroupLists];
    int64_t total_tiles[kGroupLists];
};

// Persistent grid of list w's kernel (its occupancy times the CU count).
int grouped_grid(int w, int num_cus);
// Threads per workgroup of list w's kernel.
int grouped_block(int w);
// Workgroups of one list's kernel that can hold a tile at once, for sizing
// the grid of a small list.
void *grouped_kernel_ptr(int w);
// Sorts each list by decreasing K, assigns tile ranges and writes header and
// problems; returns the total tile count, or -1 past `capacity` problems.
int64_t build_group(const std::vector<GroupProblem> (&lists)[kGroupLists], GroupHeader *hdr,
                    GroupProblem *out, int capacity);

}  // namespace blasdsp
