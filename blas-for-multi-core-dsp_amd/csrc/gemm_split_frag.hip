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
// step a workgroup needs 72 fragments (op(A): 16 row blocks x 3 pieces,
// op(B): 8 x 3); they go global -> LDS by direct-to-LDS buffer loads
// (buffer_load_dwordx4 ... lds: lane l's 16 bytes land at the LDS base +
// 16 l), no staging registers, into two LDS stages (144 KiB), one step
// ahead.  Each step a wave reads its 36 fragments and issues 192 MFMAs.
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
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kFrag = 512;      // 16-bit elements per 16 x 32 fragment (1 KiB)
constexpr int kTileRB = 16;     // op(A) row blocks of a 256-row tile
constexpr int kChunkSteps = 6;  // 32-deep K steps per accumulation chunk (192)

// The forms (scheme ids of gemm_split.hip).  NP pieces per operand; products
// (piece of op(A) a[t], piece of op(B) b[t]), smallest first; NREAD: the
// products whose pieces are read first-use first (all pieces are read by
// then).  CHUNK: two-level accumulation (each 192-deep chunk from zero into
// `part`, then added into `acc`).  FP16: scaled FP16 pieces, hi*hi into acc
// and the cross terms (carrying 2^11) into part.
template <int F>
struct FragForm;
template <>
struct FragForm<0> {  // BF16, 6 products
    static constexpr int NP = 3, NT = 6, NREAD = 3;
    static constexpr bool FP16 = false, CHUNK = true;
    static constexpr int a[6] = {1, 0, 2, 0, 1, 0};  // mm, hl, lh, hm, mh, hh
    static constexpr int b[6] = {1, 2, 0, 1, 0, 0};
};
template <>
struct FragForm<1> {  // BF16, 3 products
    static constexpr int NP = 2, NT = 3, NREAD = 2;
    static constexpr bool FP16 = false, CHUNK = true;
    static constexpr int a[3] = {0, 1, 0};  // hl, lh, hh
    static constexpr int b[3] = {1, 0, 0};
};
template <>
struct FragForm<2> {  // scaled FP16, 3 products
    static constexpr int NP = 2, NT = 3, NREAD = 2;
    static constexpr bool FP16 = true, CHUNK = false;
    static constexpr int a[3] = {0, 1, 0};  // hl, lh (part), hh (acc)
    static constexpr int b[3] = {1, 0, 0};
};

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
    __shared__ __attribute__((aligned(1024))) uint16_t lds0[St::FRAGS * kFrag];
    __shared__ __attribute__((aligned(1024))) uint16_t lds1[St::FRAGS * kFrag];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: scalar addressing
    const int lr = lane & 15, lg = lane >> 4;

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
    auto stage = [&](auto s_c) -> uint16_t * { return decltype(s_c)::value == 0 ? lds0 : lds1; };
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

    // A wave owns 128 x 64 of the 256 x 128 tile: op(A) row blocks
    // 8 (wave / 2) .. + 7, op(B) blocks 4 (wave % 2) .. + 3.
    f4 acc[8][4], part[8][4];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            acc[i][j] = f4{0, 0, 0, 0};
            part[i][j] = f4{0, 0, 0, 0};
        }

    // Step kt on buffer B = kt % 2: wait for this wave's loads of stage kt,
    // barrier (every wave's loads of stage kt have landed; every wave is done
    // with the other buffer), load stage kt + 1 into the other buffer, read
    // this wave's 12 NP fragments and issue its 32 NT MFMAs.  Past the last
    // stage a step reloads the last one (the same wait every step); that
    // buffer is never read.
    auto step = [&](int kt, auto b_c) {
        constexpr int BB = decltype(b_c)::value;
        __builtin_amdgcn_sched_barrier(0);
        wait_vm_lgkm<0, 0>();
        barrier();
        __builtin_amdgcn_sched_barrier(0);
        issue(kt + 1, std::integral_constant<int, 1 - BB>());
        const uint16_t *base = stage(b_c);
        u32x4 fa[NP][8], fb[NP][4];
        // Fragments in the order the products first use them (six products:
        // mid A and B, then hi A and lo B, then lo A and hi B; three: hi A and
        // lo B, then lo A and hi B).
#pragma unroll
        for (int t = 0; t < Fo::NREAD; ++t) {
            const int pa = Fo::a[t], pb = Fo::b[t];
#pragma unroll
            for (int i = 0; i < 8; ++i)
                fa[pa][i] =
                    *reinterpret_cast<const u32x4 *>(base + (((wave >> 1) * 8 + i) * NP + pa) * kFrag + lane * 8);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                fb[pb][j] = *reinterpret_cast<const u32x4 *>(base + (kStageA + ((wave & 1) * 4 + j) * NP + pb) * kFrag +
                                                             lane * 8);
        }
#pragma unroll
        for (int t = 0; t < Fo::NT; ++t)
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (!Fo::FP16 || t < Fo::NT - 1)
                        part[i][j] = mfma16<Fo::FP16>(fa[Fo::a[t]][i], fb[Fo::b[t]][j], part[i][j]);
                    else
                        acc[i][j] = mfma16<Fo::FP16>(fa[Fo::a[t]][i], fb[Fo::b[t]][j], acc[i][j]);
                }
        // Order within the step: the stage loads; six products: the first
        // two products' 24 LDS reads, the first product's 32 MFMAs (waiting
        // for its own 12 reads), the third product's 12 reads, the other 160
        // MFMAs; three products: the first product's 12 reads, its 32 MFMAs,
        // the second's 12 reads, the other 64 MFMAs.
        __builtin_amdgcn_sched_group_barrier(0x020, kLoadsPerWave, 0);
        if constexpr (Fo::NT == 6) {
            __builtin_amdgcn_sched_group_barrier(0x100, 24, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 32, 0);
            __builtin_amdgcn_sched_group_barrier(0x100, 12, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 160, 0);
        } else {
            __builtin_amdgcn_sched_group_barrier(0x100, 12, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 32, 0);
            __builtin_amdgcn_sched_group_barrier(0x100, 12, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 32 * (Fo::NT - 1), 0);
        }
        __builtin_amdgcn_sched_barrier(0);
    };
    auto flush = [&]() {
        if constexpr (!Fo::CHUNK) return;
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[i][j] += part[i][j];
                part[i][j] = f4{0, 0, 0, 0};
            }
    };
    using B0 = std::integral_constant<int, 0>;
    using B1 = std::integral_constant<int, 1>;

    issue(0, B0());
    // Whole chunks of 6 steps (192-deep K) run unrolled, then the chunk is
    // added into the running sum; the last, partial chunk runs the same code
    // with each step guarded.
    int kt = 0;
    for (; kt + kChunkSteps <= nk; kt += kChunkSteps) {
        step(kt + 0, B0());
        step(kt + 1, B1());
        step(kt + 2, B0());
        step(kt + 3, B1());
        step(kt + 4, B0());
        step(kt + 5, B1());
        flush();
    }
    if (kt < nk) {
        step(kt + 0, B0());
        if (kt + 1 < nk) step(kt + 1, B1());
        if (kt + 2 < nk) step(kt + 2, B0());
        if (kt + 3 < nk) step(kt + 3, B1());
        if (kt + 4 < nk) step(kt + 4, B0());
        flush();
    }
    // no load into this workgroup's LDS may be in flight when it exits
    wait_vm_lgkm<0, 15>();

    const int64_t m0 = (int64_t)tm * 256 + (wave >> 1) * 128, n0 = (int64_t)tn * 128 + (wave & 1) * 64;
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t col = n0 + j * 16 + lr;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = m0 + i * 16 + 4 * lg + r;
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
