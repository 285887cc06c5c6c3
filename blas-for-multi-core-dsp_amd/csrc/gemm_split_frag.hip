// gemm_split_frag.hip — the six-product split-BF16 FP32 GEMM on
// fragment-tiled pieces.
//
// No counterpart in the reference (an added precision form).  Each FP32
// operand x is split into three round-to-nearest BF16 pieces x = hi + mid +
// lo (residual ~2^-24 |x|) and C = alpha (sum of the six products hh + hm +
// mh + hl + lh + mm) + beta C, accumulated in FP32 in two levels (each
// 192-deep K chunk from zero, then added into the running sum).  Piece
// products are exact in FP32, so the error comes from the dropped products
// (~2^-24), the split residual and FP32 accumulation.
//
// Fragment-tiled pieces.  The pre-pass writes the pieces in the order the
// MFMA reads them: for v_mfma_f32_16x16x32_bf16 lane l of a wave supplies
// row l % 16 and k = 8 (l / 16) .. 8 (l / 16) + 7 of a 16 x 32 block, so a
// block of one piece is stored as 64 lane records of 16 bytes (1 KiB), and
// the blocks are ordered [k block][row block][piece]:
//   element (r, k) of piece p -> (((k / 32) RB + r / 16) 3 + p) 512 +
//   ((r % 16) + 16 ((k % 32) / 8)) 8 + k % 8,   RB = Rp / 16.
// All that a tile needs for one 32-deep K step (its row blocks x 3 pieces)
// is then one contiguous run: the loads spread over every memory channel,
// where blocks a power-of-two stride apart piled onto a few.  One fragment
// is one conflict-free 16-byte-per-lane read from LDS.
//
// GEMM kernel.  A 256 x 128 tile of C per workgroup, 4 waves of 128 x 64
// (chunk and running accumulators, 256 registers, in AGPRs).  Per 32-deep K
// step a workgroup needs 72 fragments with three pieces (op(A): 16 row
// blocks x 3 pieces, op(B): 8 x 3), 48 with two; they go global -> LDS by
// direct-to-LDS buffer loads (buffer_load_dwordx4 ... lds: lane l's 16
// bytes land at the LDS base + 16 l), no staging registers, into a ring of
// LDS stages (two of 72 KiB or three of 48 KiB).  Each step a wave reads
// its 36 (24) fragments and issues 192 (96) MFMAs' worth of products.
// What bounds it, measured at 8192^3 with timing-only copies (six products,
// constant operands): the global -> LDS loads alone take 3.2 ms, the MFMAs
// alone with the loads 3.3 ms, the whole kernel 5.0; the loads hit L2 79 %
// of the time, the compulsory rate of the 8 x 4 tiles an XCD runs at once.
// Measured at 8192^3 on MI355X (each change bit-identical in the tests):
// row-major pieces with register staging and 128 x 128 tiles 8.46 ms per
// call; fragment tiling, direct-to-LDS loads and a three-stage ring 6.96;
// fixed code positions for the chunk flush and the waits 6.10; [k][row]
// block order 5.90; 256 x 128 tiles 5.45.  Register staging with ds_write
// in place of the direct-to-LDS loads ran 6.27 (128 x 128 tiles).
#include <type_traits>

#include "internal.h"

namespace blasdsp {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kFrag = 512;      // 16-bit elements per 16 x 32 fragment (1 KiB)
constexpr int kTileRB = 16;     // op(A) row blocks of a 256-row tile
constexpr int kChunkSteps = 6;  // 32-deep K steps per accumulation chunk (192)

// The forms (scheme ids of gemm_split.hip).  NP pieces per operand; products
// (piece of op(A) a[t], piece of op(B) b[t]) in issue order.  CHUNK:
// two-level accumulation (each 192-deep chunk from zero into `part`, then
// added into `acc`).  FP16: scaled FP16 pieces, hi*hi into acc and the cross
// terms (carrying 2^11) into part.
//
// Fragment schedule of one K step (the pieces' registers are reused across
// steps): the HEAD pieces (op(A) piece when hb = 0, op(B) when 1) are read
// from LDS at the step's start; the others were read during the previous
// step, each right after that step's last product using the register it
// goes into (pf[t]: the piece read after product t, -1 none; pb[t]: op(B)).
// The first product of a step uses only such prefetched pieces, so its
// MFMAs start with no wait on LDS; the workgroup barrier that makes the next
// stage readable sits after product BAR, before the first prefetch.
// NSTAGE: LDS stages in the ring (72 KiB each with three pieces, 48 KiB
// with two).  MF32: the products run on 32 x 32 x 16 MFMAs instead of
// 16 x 16 x 32 ones (same fragments, other lane addresses).  Measured at
// 8192^3 on MI355X, all versions in one process on the same pieces of
// random operands (medians of 5): six BF16 products 5.32 -> 5.01 ms with this
// schedule (5.12 with 32 x 32 MFMAs); three BF16 products 2.92 -> 2.72 (a
// third stage and this schedule; 2.76 with 32 x 32); scaled FP16 2.86 ->
// 2.80, and 2.71 with 32 x 32.  Four tile rows per XCD group in place of
// eight (fewer bytes per step from beyond L2) measured within noise.
template <int F>
struct FragForm;
template <>
struct FragForm<0> {  // BF16, 6 products
    static constexpr int NP = 3, NT = 6, NSTAGE = 2;
    static constexpr bool MF32 = false;
    static constexpr bool FP16 = false, CHUNK = true;
    static constexpr int a[6] = {1, 1, 2, 0, 0, 0};  // mm, mh, lh, hm, hl, hh
    static constexpr int b[6] = {1, 0, 0, 1, 2, 0};
    static constexpr int NH = 2, hb[2] = {1, 0}, hp[2] = {0, 0};  // op(B) hi, op(A) hi
    static constexpr int BAR = 1;
    static constexpr int pf[6] = {-1, 1, 2, 1, 2, -1};  // A mid, A lo, B mid, B lo
    static constexpr int pb[6] = {0, 0, 0, 1, 1, 0};
};
template <>
struct FragForm<1> {  // BF16, 3 products
    static constexpr int NP = 2, NT = 3, NSTAGE = 3;
    static constexpr bool MF32 = false;
    static constexpr bool FP16 = false, CHUNK = true;
    static constexpr int a[3] = {1, 0, 0};  // lh, hh, hl
    static constexpr int b[3] = {0, 0, 1};
    static constexpr int NH = 2, hb[2] = {0, 1}, hp[2] = {0, 1};  // op(A) hi, op(B) lo
    static constexpr int BAR = 0;
    static constexpr int pf[3] = {1, 0, -1};  // A lo, B hi
    static constexpr int pb[3] = {0, 1, 0};
};
template <>
struct FragForm<2> {  // scaled FP16, 3 products
    static constexpr int NP = 2, NT = 3, NSTAGE = 3;
    static constexpr bool MF32 = true;
    static constexpr bool FP16 = true, CHUNK = false;
    static constexpr int a[3] = {1, 0, 0};  // lh (part), hh (acc), hl (part)
    static constexpr int b[3] = {0, 0, 1};
    static constexpr int NH = 2, hb[2] = {0, 1}, hp[2] = {0, 1};
    static constexpr int BAR = 0;
    static constexpr int pf[3] = {1, 0, -1};
    static constexpr int pb[3] = {0, 1, 0};
};

// Each piece of each operand is read exactly once per step: at the head or
// as one product's prefetch, and that only after its last use in the step.
template <int F>
constexpr bool schedule_ok()
{
    using Fo = FragForm<F>;
    for (int side = 0; side < 2; ++side)
        for (int p = 0; p < Fo::NP; ++p) {
            int reads = 0, at = -1;  // at: product after which it is read (-1: head)
            for (int h = 0; h < Fo::NH; ++h)
                if (Fo::hb[h] == side && Fo::hp[h] == p) ++reads;
            for (int t = 0; t < Fo::NT; ++t)
                if (Fo::pf[t] == p && Fo::pb[t] == side) {
                    ++reads;
                    at = t;
                    if (t < Fo::BAR) return false;  // before the barrier
                }
            if (reads != 1) return false;
            for (int t = at + 1; t < Fo::NT && at >= 0; ++t)
                if ((side == 0 ? Fo::a[t] : Fo::b[t]) == p) return false;  // used after its prefetch
            // a head piece is never used by product 0 (its read is the step's first)
            if (at < 0 && (side == 0 ? Fo::a[0] : Fo::b[0]) == p) return false;
        }
    return true;
}
static_assert(schedule_ok<0>() && schedule_ok<1>() && schedule_ok<2>(), "fragment schedule");

// Buffer of step J of a chunk with NS stages.
template <int J, int NS>
using BufOf = std::integral_constant<int, J % NS>;

template <int F>
struct Stage {
    static constexpr int NP = FragForm<F>::NP;
    static constexpr int A = kTileRB * NP;      // op(A) fragments per stage (16 KiB per piece)
    static constexpr int B = kTileRB / 2 * NP;  // op(B) fragments per stage (8 KiB per piece)
    static constexpr int FRAGS = A + B;
    static constexpr int LOADS = FRAGS / 4;     // per wave
};

__device__ __forceinline__ uint32_t bf16_bits(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ float bf16_val(uint32_t h) { return __uint_as_float(h << 16); }

// Pieces of one FP32 value.  BF16: x = p0 + p1 (+ p2) + r, each piece the
// round-to-nearest BF16 of what is left.  FP16 (scaled): with x' = x * scale
// (a power of two, exact), hi = fp16(x'), lo = fp16((x' - hi) * 2^11); x' - hi
// is exact in FP32 and the pair carries 22 significant bits.
template <int NP, bool FP16>
__device__ __forceinline__ void pieces(float x, float scale, uint32_t (&p)[NP])
{
    if constexpr (FP16) {
        const float xs = x * scale;
        const _Float16 hi = (_Float16)xs;
        const float r = xs - (float)hi;
        p[0] = __builtin_bit_cast(uint16_t, hi);
        p[NP - 1] = __builtin_bit_cast(uint16_t, (_Float16)(r * 2048.0f));
    } else {
        float r = x;
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            p[q] = bf16_bits(r);
            r = r - bf16_val(p[q]);
        }
    }
}

// One 64 (rows) x 64 (k) tile of op(X) (R x K) into the NP fragment-tiled
// planes (Rp x Kp each, zero outside R x K).  256 threads, two lane records
// each: record e = tid + 256 i is lane e % 64 of fragment e / 64, fragment
// (row block f / 2, k block f % 2) of the tile.  KC: X stored [R][K] (k
// contiguous), each record's 8 values read straight; else X stored [K][R],
// read coalesced along r into an LDS tile first.  VEC: 16-byte loads allowed.
// FP16: row r is scaled by scale[r] (power of two) before the split.
template <int NP, bool FP16, bool KC, bool VEC>
__global__ void __launch_bounds__(256) split_frag_kernel(const float *__restrict__ X, int64_t ld, int64_t R,
                                                        int64_t K, uint16_t *__restrict__ planes, int64_t RB,
                                                        const float *__restrict__ scale)
{
    __shared__ float tile[64][65];  // [r][k]
    const int tid = threadIdx.x;
    const int64_t k0 = (int64_t)blockIdx.x * 64, r0 = (int64_t)blockIdx.y * 64;
    if (!KC) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int e = tid + i * 256;
            const int kk = e / 16, rq = (e % 16) * 4;
            const int64_t gk = k0 + kk, gr = r0 + rq;
            f4v v = {0.0f, 0.0f, 0.0f, 0.0f};
            if (gk < K) {
                if (VEC && gr + 3 < R) {
                    v = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(X + gk * ld + gr));
                } else {
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        if (gr + c < R) v[c] = X[gk * ld + gr + c];
                }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) tile[rq + c][kk] = v[c];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int e = tid + i * 256, f = e >> 6, lane = e & 63;
        const int rl = (f >> 1) * 16 + (lane & 15), kl = (f & 1) * 32 + (lane >> 4) * 8;
        const int64_t gr = r0 + rl, gk = k0 + kl;
        float x[8];
        if (KC) {
#pragma unroll
            for (int j = 0; j < 8; ++j) x[j] = 0.0f;
            if (gr < R) {
                if (VEC && gk + 7 < K) {
                    const f4v a = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(X + gr * ld + gk));
                    const f4v b = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(X + gr * ld + gk + 4));
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        x[j] = a[j];
                        x[4 + j] = b[j];
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < 8; ++j)
                        if (gk + j < K) x[j] = X[gr * ld + gk + j];
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) x[j] = tile[rl][kl + j];
        }
        const float sc = FP16 ? scale[gr] : 1.0f;  // scales cover the padded rows
        uint32_t p[NP][8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            uint32_t pj[NP];
            pieces<NP, FP16>(x[j], sc, pj);
#pragma unroll
            for (int q = 0; q < NP; ++q) p[q][j] = pj[q];
        }
        const int64_t frag = (((k0 >> 5) + (f & 1)) * RB + (r0 >> 4) + (f >> 1)) * NP;
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            u32x4 w;
#pragma unroll
            for (int j = 0; j < 4; ++j) w[j] = p[q][2 * j] | (p[q][2 * j + 1] << 16);
            *reinterpret_cast<u32x4 *>(planes + ((frag + q) * 64 + lane) * 8) = w;
        }
    }
}

// 16 x 16 x 32 product: lane l supplies row (column) l % 16 and k = 8 (l /
// 16) .. + 7 of each operand (one fragment record); register r of lane l
// holds output row 4 (l / 16) + r, column l % 16.
template <bool FP16>
__device__ __forceinline__ f4 mfma16(u32x4 a, u32x4 b, f4 c)
{
    if constexpr (FP16)
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b),
                                                      c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                       c, 0, 0, 0);
}

// 32 x 32 x 16 product: lane l supplies row (column) l % 32 and k = 8 (l /
// 32) .. + 7 of each operand; register r of lane l holds output row 8 (r /
// 4) + 4 (l / 32) + r % 4, column l % 32.
template <bool FP16>
__device__ __forceinline__ f16v mfma32(u32x4 a, u32x4 b, f16v c)
{
    if constexpr (FP16)
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b),
                                                      c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                       c, 0, 0, 0);
}

// s_waitcnt vmcnt(VM) lgkmcnt(LGKM) through the builtin, so the compiler's
// own wait insertion sees it (gfx9 encoding: vmcnt bits 3:0 and 15:14,
// expcnt 6:4 left at its maximum, lgkmcnt 11:8).
template <int VM, int LGKM>
__device__ __forceinline__ void wait_vm_lgkm()
{
    static_assert(VM >= 0 && VM < 64 && LGKM >= 0 && LGKM < 16, "counter range");
    __builtin_amdgcn_s_waitcnt((VM & 15) | ((VM >> 4) << 14) | (7 << 4) | (LGKM << 8));
}

// Workgroup barrier with the compiler kept from moving memory accesses across
// it; the caller has already waited for the accesses that must complete.
__device__ __forceinline__ void barrier()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

template <int F>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
frag_kernel(const uint16_t *__restrict__ Ap, const uint16_t *__restrict__ Bp, int64_t KB, int64_t a_rb,
            int64_t b_rb, int tiles_m, int tiles_n, float *__restrict__ C, int64_t ldc, int64_t M, int64_t N,
            float alpha, float beta, const float *__restrict__ a_inv, const float *__restrict__ b_inv)
{
    using Fo = FragForm<F>;
    using St = Stage<F>;
    constexpr int NP = Fo::NP;
    constexpr int kStageA = St::A, kStageB = St::B, kLoadsPerWave = St::LOADS;
    constexpr int NS = Fo::NSTAGE;
    static_assert(NS * St::FRAGS * kFrag * 2 <= 160 * 1024, "stages fit in LDS");
    __shared__ __attribute__((aligned(1024))) uint16_t lds[NS][St::FRAGS * kFrag];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar addressing

    // XCD-aware grouped tile order (as the tile kernel's).
    int pid = blockIdx.x;
    const int nwg = gridDim.x;
    if ((nwg & 7) == 0) pid = (pid & 7) * (nwg >> 3) + (pid >> 3);
    const int width = 8 * tiles_n;
    const int first_m = (pid / width) * 8;
    const int gsz = min(tiles_m - first_m, 8);
    const int tm = first_m + (pid % width) % gsz, tn = (pid % width) / gsz;

    // Loads: a stage is op(A)'s run (16 row blocks x NP pieces, 16 NP KiB)
    // then op(B)'s (8 x NP); wave w loads fragments L w .. L w + L - 1 (L =
    // 6 NP), each a 1 KiB buffer load into LDS (lane: 16 bytes; scalar
    // offset: the fragment).  Buffer loads into LDS are counted by vmcnt
    // alone, so the compiler's waits for LDS reads do not include them.
    const uint16_t *runA = Ap + (int64_t)tm * kTileRB * NP * kFrag;
    const uint16_t *runB = Bp + (int64_t)tn * (kTileRB / 2) * NP * kFrag;
    const int64_t stepA = a_rb * NP * kFrag, stepB = b_rb * NP * kFrag;  // elements from one K block to the next
    const int voff = lane * 16;
    const int nk = (int)KB;  // 32-deep K steps
    auto stage = [&](auto s_c) -> uint16_t * { return lds[decltype(s_c)::value]; };
    auto issue = [&](int kb, auto s_c) {
        uint16_t *base = stage(s_c);
        const int64_t k = min(kb, nk - 1);
        const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<uint16_t *>(runA + k * stepA), (short)0, kStageA * kFrag * 2, 0x00020000);
        const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<uint16_t *>(runB + k * stepB), (short)0, kStageB * kFrag * 2, 0x00020000);
#pragma unroll
        for (int q = 0; q < kLoadsPerWave; ++q) {
            const int f = wave * kLoadsPerWave + q;  // wave-uniform
            const bool isb = f >= kStageA;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(isb ? rb : ra,
                                                     (__attribute__((address_space(3))) void *)(base + f * kFrag), 16,
                                                     voff, (isb ? f - kStageA : f) * (kFrag * 2), 0, 0);
        }
    };

    // A wave owns 128 x 64 of the 256 x 128 tile: op(A) rows 128 (wave / 2)
    // .., op(B) columns 64 (wave % 2) .., as TI x TJ MFMA blocks (16 x 16
    // blocks, or 32 x 32 with MF32).
    constexpr bool MF32 = Fo::MF32;
    constexpr int TI = MF32 ? 4 : 8, TJ = MF32 ? 2 : 4, RR = MF32 ? 16 : 4;
    using acc_t = typename std::conditional<MF32, f16v, f4>::type;
    acc_t acc[TI][TJ], part[TI][TJ];
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
#pragma unroll
            for (int r = 0; r < RR; ++r) {
                acc[i][j][r] = 0.0f;
                part[i][j][r] = 0.0f;
            }

    // This wave's fragments of each piece (8 of op(A), 4 of op(B)); they
    // persist across steps.  16 x 16 blocks: fa[p][i] is 16-row fragment i
    // as stored (lane l reads record l).  32 x 32 blocks: fa[p][2 i + h]
    // (fb[p][2 j + h]) is the 32 x 16 operand of block i (j) for k = 16 h ..
    // 16 h + 15 of the step: lane l's 16 bytes (row l % 32, k = 16 h + 8 (l /
    // 32) .. + 7) are record (l % 16) + 16 (2 h + l / 32) of the 16-row
    // fragment holding that row, so each 8-lane group still reads 128
    // contiguous bytes.
    u32x4 fa[NP][8], fb[NP][4];
    const int lrec = MF32 ? (((lane & 31) >> 4) * NP * kFrag) + ((lane & 15) + 16 * (lane >> 5)) * 8 : lane * 8;
    auto read_piece = [&](const uint16_t *base, int side, int p) {
        if (side == 0) {
#pragma unroll
            for (int f = 0; f < 8; ++f) {
                const int rb = MF32 ? 2 * (f >> 1) : f, ko = MF32 ? (f & 1) * 256 : 0;
                fa[p][f] = *reinterpret_cast<const u32x4 *>(base + (((wave >> 1) * 8 + rb) * NP + p) * kFrag + lrec + ko);
            }
        } else {
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                const int cb = MF32 ? 2 * (f >> 1) : f, ko = MF32 ? (f & 1) * 256 : 0;
                fb[p][f] = *reinterpret_cast<const u32x4 *>(base + (kStageA + ((wave & 1) * 4 + cb) * NP + p) * kFrag +
                                                            lrec + ko);
            }
        }
    };
    auto product = [&](int t) {
        const int pa = Fo::a[t], pb = Fo::b[t];
        const bool to_acc = Fo::FP16 && pa == 0 && pb == 0;
        if constexpr (MF32) {
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int i = 0; i < TI; ++i)
#pragma unroll
                    for (int j = 0; j < TJ; ++j) {
                        if (to_acc) acc[i][j] = mfma32<Fo::FP16>(fa[pa][2 * i + h], fb[pb][2 * j + h], acc[i][j]);
                        else part[i][j] = mfma32<Fo::FP16>(fa[pa][2 * i + h], fb[pb][2 * j + h], part[i][j]);
                    }
        } else {
#pragma unroll
            for (int i = 0; i < TI; ++i)
#pragma unroll
                for (int j = 0; j < TJ; ++j) {
                    if (to_acc) acc[i][j] = mfma16<Fo::FP16>(fa[pa][i], fb[pb][j], acc[i][j]);
                    else part[i][j] = mfma16<Fo::FP16>(fa[pa][i], fb[pb][j], part[i][j]);
                }
        }
    };
    constexpr int kMfmaPerProduct = TI * TJ * (MF32 ? 2 : 1);
    // Step kt on buffer BB = kt % 2 (the prefetched pieces of stage kt are in
    // registers): read the head pieces, issue products 0 .. BAR; wait for
    // this wave's loads of stage kt + 1 and its reads of stage kt, barrier
    // (stage kt + 1 has landed everywhere and no wave reads buffer BB any
    // more), load stage kt + 2 into buffer BB, then the remaining products,
    // each followed by the prefetch from stage kt + 1 of the piece it freed.
    // Past the last stage a step reloads the last one (the same wait every
    // step); those pieces are never used.
    auto step = [&](int kt, auto b_c) {
        constexpr int BB = decltype(b_c)::value;
        const uint16_t *cur = stage(b_c);
        const uint16_t *nxt = stage(std::integral_constant<int, (BB + 1) % NS>());
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int h = 0; h < Fo::NH; ++h) read_piece(cur, Fo::hb[h], Fo::hp[h]);
#pragma unroll
        for (int t = 0; t <= Fo::BAR; ++t) product(t);
        __builtin_amdgcn_sched_group_barrier(0x100, Fo::NH == 2 ? 12 : 0, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, kMfmaPerProduct * (Fo::BAR + 1), 0);
        __builtin_amdgcn_sched_barrier(0);
        wait_vm_lgkm<(NS - 2) * kLoadsPerWave, 0>();  // stage kt + 1 (later stages may be in flight)
        barrier();
        __builtin_amdgcn_sched_barrier(0);
        issue(kt + NS, b_c);
#pragma unroll
        for (int t = Fo::BAR + 1; t < Fo::NT; ++t) {
            if (Fo::pf[t - 1] >= 0) read_piece(nxt, Fo::pb[t - 1], Fo::pf[t - 1]);
            product(t);
        }
        if (Fo::pf[Fo::NT - 1] >= 0) read_piece(nxt, Fo::pb[Fo::NT - 1], Fo::pf[Fo::NT - 1]);
        // the stage loads, then per product its prefetched piece's reads
        // (from the product before it) and its 32 MFMAs
        __builtin_amdgcn_sched_group_barrier(0x020, kLoadsPerWave, 0);
#pragma unroll
        for (int t = Fo::BAR + 1; t < Fo::NT; ++t) {
            if (Fo::pf[t - 1] >= 0) {
                if (Fo::pb[t - 1]) __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
                else __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);
            }
            __builtin_amdgcn_sched_group_barrier(0x008, kMfmaPerProduct, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
    };
    auto flush = [&]() {
        if constexpr (!Fo::CHUNK) return;
#pragma unroll
        for (int i = 0; i < TI; ++i)
#pragma unroll
            for (int j = 0; j < TJ; ++j)
#pragma unroll
                for (int r = 0; r < RR; ++r) {
                    acc[i][j][r] += part[i][j][r];
                    part[i][j][r] = 0.0f;
                }
    };
    // step j of a chunk (chunks start at multiples of 6 steps, a multiple
    // of the stage count) uses buffer j % NS
    static_assert(kChunkSteps % NS == 0, "chunks start on buffer 0");
    using B0 = std::integral_constant<int, 0>;
    using B1 = std::integral_constant<int, 1>;

    // Stages 0 .. NS - 1 issued, stage 0 landed everywhere, and step 0's
    // prefetched pieces read from it.
    issue(0, B0());
    if constexpr (NS == 3) issue(1, B1());
    wait_vm_lgkm<(NS - 2) * kLoadsPerWave, 0>();
    barrier();
    issue(NS - 1, std::integral_constant<int, NS - 1>());
#pragma unroll
    for (int t = 0; t < Fo::NT; ++t)
        if (Fo::pf[t] >= 0) read_piece(lds[0], Fo::pb[t], Fo::pf[t]);
    // Whole chunks of 6 steps (192-deep K) run unrolled, then the chunk is
    // added into the running sum; the last, partial chunk runs the same code
    // with each step guarded.
    int kt = 0;
    for (; kt + kChunkSteps <= nk; kt += kChunkSteps) {
        step(kt + 0, BufOf<0, NS>());
        step(kt + 1, BufOf<1, NS>());
        step(kt + 2, BufOf<2, NS>());
        step(kt + 3, BufOf<3, NS>());
        step(kt + 4, BufOf<4, NS>());
        step(kt + 5, BufOf<5, NS>());
        flush();
    }
    if (kt < nk) {
        step(kt + 0, BufOf<0, NS>());
        if (kt + 1 < nk) step(kt + 1, BufOf<1, NS>());
        if (kt + 2 < nk) step(kt + 2, BufOf<2, NS>());
        if (kt + 3 < nk) step(kt + 3, BufOf<3, NS>());
        if (kt + 4 < nk) step(kt + 4, BufOf<4, NS>());
        flush();
    }
    // no load into this workgroup's LDS may be in flight when it exits
    wait_vm_lgkm<0, 15>();

    const int64_t m0 = (int64_t)tm * 256 + (wave >> 1) * 128, n0 = (int64_t)tn * 128 + (wave & 1) * 64;
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) {
            const int64_t col = MF32 ? n0 + j * 32 + (lane & 31) : n0 + j * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < RR; ++r) {
                const int64_t row = MF32 ? m0 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)
                                         : m0 + i * 16 + 4 * (lane >> 4) + r;
                if (row < M && col < N) {
                    float *p = C + row * ldc + col;
                    float v;
                    if constexpr (Fo::FP16)
                        v = alpha * ((acc[i][j][r] + 0x1.0p-11f * part[i][j][r]) * a_inv[row] * b_inv[col]);
                    else
                        v = alpha * acc[i][j][r];
                    if (beta != 0.0f) v += beta * *p;
                    *p = v;
                }
            }
        }
}

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

template <int NP, bool FP16>
hipError_t split_frag(const float *X, int64_t ld, bool kc, int64_t R, int64_t K, uint16_t *planes, int64_t Rp,
                      int64_t Kp, const float *scale, hipStream_t s)
{
    if (Rp / 64 > 65535 || Kp / 64 > 0x7fffffffLL) return hipErrorInvalidValue;  // grid limits
    const dim3 grid((unsigned)(Kp / 64), (unsigned)(Rp / 64));
    const bool vec = ((uintptr_t)X % 16) == 0 && ld % 4 == 0;
    const int64_t RB = Rp / 16;
    if (kc && vec)
        hipLaunchKernelGGL((split_frag_kernel<NP, FP16, true, true>), grid, dim3(256), 0, s, X, ld, R, K, planes, RB, scale);
    else if (kc)
        hipLaunchKernelGGL((split_frag_kernel<NP, FP16, true, false>), grid, dim3(256), 0, s, X, ld, R, K, planes, RB, scale);
    else if (vec)
        hipLaunchKernelGGL((split_frag_kernel<NP, FP16, false, true>), grid, dim3(256), 0, s, X, ld, R, K, planes, RB, scale);
    else
        hipLaunchKernelGGL((split_frag_kernel<NP, FP16, false, false>), grid, dim3(256), 0, s, X, ld, R, K, planes, RB, scale);
    return hipGetLastError();
}

template <int F>
hipError_t run_frag(const Gemm<float> &g, uint16_t *Ap, uint16_t *Bp, int64_t Mp, int64_t Np, int64_t Kp,
                    const float *a_scale, const float *a_inv, const float *b_scale, const float *b_inv, hipStream_t s)
{
    using Fo = FragForm<F>;
    // op(A) rows: k contiguous unless A is transposed; op(B)^T rows: k
    // contiguous when B is transposed.
    hipError_t e = split_frag<Fo::NP, Fo::FP16>(g.A, g.lda, !g.ta, g.M, g.K, Ap, Mp, Kp, a_scale, s);
    if (e == hipSuccess) e = split_frag<Fo::NP, Fo::FP16>(g.B, g.ldb, g.tb, g.N, g.K, Bp, Np, Kp, b_scale, s);
    if (e != hipSuccess) return e;
    const int64_t tiles_m = Mp / 256, tiles_n = Np / 128, nwg = tiles_m * tiles_n;
    if (nwg > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(frag_kernel<F>, dim3((unsigned)nwg), dim3(256), 0, s, Ap, Bp, Kp / 32, Mp / 16, Np / 16,
                       (int)tiles_m, (int)tiles_n, g.C, g.ldc, g.M, g.N, g.alpha, g.beta, a_inv, b_inv);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_frag_gemm(const Gemm<float> &g, int scheme, uint16_t *Ap, uint16_t *Bp, int64_t Mp, int64_t Np,
                            int64_t Kp, const float *a_scale, const float *a_inv, const float *b_scale,
                            const float *b_inv, hipStream_t s)
{
    if (Mp % 256 || Np % 128 || Kp % 64) return hipErrorInvalidValue;
    switch (scheme) {
    case 0: return run_frag<0>(g, Ap, Bp, Mp, Np, Kp, nullptr, nullptr, nullptr, nullptr, s);
    case 1: return run_frag<1>(g, Ap, Bp, Mp, Np, Kp, nullptr, nullptr, nullptr, nullptr, s);
    case 2:
        if (!a_scale || !a_inv || !b_scale || !b_inv) return hipErrorInvalidValue;
        return run_frag<2>(g, Ap, Bp, Mp, Np, Kp, a_scale, a_inv, b_scale, b_inv, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace blasdsp
