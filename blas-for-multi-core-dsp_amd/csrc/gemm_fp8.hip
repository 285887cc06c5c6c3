// gemm_fp8.hip — FP32 GEMM through the FP8 (E4M3) matrix cores with one
// power-of-two scale per 128-element block of every row of op(A) and every
// column of op(B).  No counterpart in the reference (whose sgemm is FP32 FMA
// throughout, sgemm_release_V1.3.2/device_code/sgemm_kernel.c:1384): a
// reduced-precision form offered beside the plain and split FP32 kernels, with
// its error reported rather than bounded.
//
// Two passes:
//   quant_kc_kernel / quant_mc_kernel  read op(A) rows and op(B) columns (K-
//       contiguous or not, whichever way they are stored), pick each 128-block's
//       exponent e (the smallest with amax 2^-e <= 448), and write the E4M3
//       codes of x 2^-e (round to nearest even) K-contiguously, zero-padded to
//       whole tiles in every dimension, plus the biased exponents 127 + e
//       (E8M0 scale bytes);
//   fp8_big_kernel  K steps of 128 = one v_mfma_scale_f32_16x16x128_f8f6f4
//       depth, so each step's scales are one byte per row / column, applied
//       by the matrix core itself; the quantized operands and each wave's
//       scale words go global -> LDS by direct-to-LDS loads, the operands into
//       an XOR-swizzled layout.  128 x 128 tiles of 4 waves, two workgroups
//       per CU; 256 x 256 tiles of 8 waves, one per CU, when a launch has at
//       least four of those per CU.  At 8192^3 on MI355X the whole form
//       (quantization included) went from 0.88-0.90 ms with fp8_gemm_kernel
//       (the same tiles staged through registers into padded LDS rows) to
//       0.74-0.76 ms with direct-to-LDS loads; 256 x 128 and 256 x 256 tiles
//       of 4 waves at one workgroup per CU ran 0.97-1.0 ms.  On one box, all
//       bit-identical: the scale words through LDS in place of registers
//       (whose loads the compiler waited on at every loop entry) 0.77-0.79 ->
//       0.73-0.76 ms, 256 x 256 tiles of 8 waves 0.69-0.74.
//       fp8_gemm_kernel remains for K past the direct loads' 32-bit lane
//       offsets (8 Kp >= 2^31).
//
// MFMA layout, measured with a lane probe: A lane l holds 32 bytes of row
// l % 16 at k = 32 (l / 16) .. +31 (B the same for column l % 16, the same k),
// lane l's scale byte applies to row (column) l % 16 and k block l / 16, and
// output register r of lane l is row 4 (l / 16) + r, column l % 16.
#include <math.h>

#include <type_traits>
#include <utility>

#include "api_common.h"

namespace blasdsp {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef int i8v __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kQB = 128;         // k per scale block = MFMA depth = K step
constexpr int kTile = 128;       // output tile edge
constexpr int kRow = 144;        // LDS row stride in bytes: 128 data + 16 pad
constexpr int kRowBlock = 64;    // rows whose four scales of one k block share a word

// Smallest e with amax 2^-e <= 448, clamped to [-127, 127]; 0 for a zero block.
__device__ __forceinline__ int block_exp(float amax)
{
    if (!(amax > 0.0f)) return 0;
    int x;
    const float f = frexpf(amax, &x);
    const int e = x - 9 + (f > 0.875f ? 1 : 0);
    return e < -127 ? -127 : (e > 127 ? 127 : e);
}

// Byte offset of the scale of row r (of Rp padded rows) in k block kb: the
// four rows r0 + 16 i + l (i = 0..3) of a 64-row group are one 32-bit word,
// so a lane of the GEMM reads its four fragments' scales with one load.
__device__ __forceinline__ int64_t scale_byte(int64_t r, int64_t kb, int64_t Rp)
{
    return ((kb * (Rp / kRowBlock) + r / kRowBlock) * 16 + (r % 16)) * 4 + (r % kRowBlock) / 16;
}

typedef float f4v __attribute__((ext_vector_type(4)));

// Four E4M3 codes (OCP: bias 7, subnormal step 2^-9, 0x7F NaN) of x 2^-e,
// round to nearest even, packed little-endian.  The scaled values satisfy
// |y| <= 448 unless the block holds a non-finite value.  v_cvt_pk_fp8_f32
// converts two values per instruction; on MI355X it measured bit-identical
// to a software round-to-nearest-even encoder (the oracle's rule) on every
// FP32 value in [-448, 448], and replacing that encoder (frexp, ldexp and
// rint per element) with it made both quantizers bandwidth-bound.
// Non-finite values get the NaN code, as in the oracle.
__device__ __forceinline__ uint32_t pack4(f4v x, int e)
{
    f4v y;
#pragma unroll
    for (int i = 0; i < 4; ++i) y[i] = ldexpf(x[i], -e);
    int w = __builtin_amdgcn_cvt_pk_fp8_f32(y[0], y[1], 0, false);
    w = __builtin_amdgcn_cvt_pk_fp8_f32(y[2], y[3], w, true);
    uint32_t u = (uint32_t)w;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (!(fabsf(y[i]) < INFINITY)) u = (u & ~(0xFFu << (8 * i))) | (0x7Fu << (8 * i));
    return u;
}

// Rows of op(X) stored K-contiguously: element (r, k) at X[r ld + k].  A
// half wave takes one (row, k block), four consecutive k per lane; each wave
// walks kQuantBlocks pairs of blocks.  VEC: 16-byte loads (X 16-byte aligned
// and ld a multiple of 4).
constexpr int kQuantBlocks = 4;

template <bool VEC>
__global__ void __launch_bounds__(256) quant_kc_kernel(const float *__restrict__ X, int64_t ld, int64_t R,
                                                       int64_t K, int64_t Rp, int64_t KB,
                                                       uint8_t *__restrict__ Q, uint8_t *__restrict__ S)
{
    const int lane = threadIdx.x & 63, half = lane >> 5, hl = lane & 31;
    const int64_t wave_id = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t nblk = Rp * KB;
#pragma unroll
    for (int it = 0; it < kQuantBlocks; ++it) {
        const int64_t id = (wave_id * kQuantBlocks + it) * 2 + half;
        if (id >= nblk) break;  // uniform per half wave; no cross-half exchange below
        const int64_t r = id / KB, kb = id % KB, k = kb * kQB + 4 * hl;
        f4v x = {0.0f, 0.0f, 0.0f, 0.0f};
        if (r < R) {
            if (VEC && k + 3 < K) {
                x = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(X + r * ld + k));
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (k + e < K) x[e] = X[r * ld + k + e];
            }
        }
        float amax = fmaxf(fmaxf(fabsf(x[0]), fabsf(x[1])), fmaxf(fabsf(x[2]), fabsf(x[3])));
#pragma unroll
        for (int o = 16; o > 0; o >>= 1) amax = fmaxf(amax, __shfl_xor(amax, o));
        const int e = block_exp(amax);
        *reinterpret_cast<uint32_t *>(Q + r * (KB * kQB) + k) = pack4(x, e);
        if (hl == 0) S[scale_byte(r, kb, Rp)] = (uint8_t)(127 + e);
    }
}

// Rows of op(X) stored transposed: element (r, k) at X[k ld + r].  A
// workgroup takes 64 rows x one k block through LDS.  Loads: thread
// (r4, kq) = (t % 16, t / 16) reads rows 4 r4 .. + 3 of k rows kq + 16 i as
// one 16-byte vector (VEC: X 16-byte aligned, ld a multiple of 4); then each
// thread quantizes 32 consecutive k of one row, four lanes to a row, so a
// wave's stores are 16 runs of 128 bytes (with a lane to a row, 64 runs of
// 32 bytes: 86 against 78 us for an 8192 x 8192 operand on MI355X).
template <bool VEC>
__global__ void __launch_bounds__(256) quant_mc_kernel(const float *__restrict__ X, int64_t ld, int64_t R,
                                                       int64_t K, int64_t Rp, int64_t KB,
                                                       uint8_t *__restrict__ Q, uint8_t *__restrict__ S)
{
    __shared__ float T[64][kQB + 1];
    __shared__ float wmax[4][64];
    const int tid = threadIdx.x, r4 = tid & 15, kq = tid >> 4;
    const int64_t r0 = (int64_t)blockIdx.x * 64, kb = blockIdx.y, k0 = kb * kQB;
    const int64_t rv = r0 + 4 * r4;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int k = kq + 16 * i;
        f4v v = {0.0f, 0.0f, 0.0f, 0.0f};
        if (k0 + k < K) {
            if (VEC && rv + 3 < R) {
                v = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(X + (k0 + k) * ld + rv));
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (rv + e < R) v[e] = X[(k0 + k) * ld + rv + e];
            }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) T[4 * r4 + e][k] = v[e];
    }
    __syncthreads();
    // block maxima: thread (rl, w) takes row rl, k w * 32 .. + 31
    {
        const int rl = tid & 63, w = tid >> 6;
        float amax = 0.0f;
#pragma unroll
        for (int j = 0; j < 32; ++j) amax = fmaxf(amax, fabsf(T[rl][w * 32 + j]));
        wmax[w][rl] = amax;
    }
    __syncthreads();
    // codes: thread t takes row t / 4, k (t % 4) * 32 .. + 31, so four
    // consecutive lanes write one row's 128 bytes
    const int rq = tid >> 2, seg = tid & 3;
    const int64_t r = r0 + rq;
    const float amax = fmaxf(fmaxf(wmax[0][rq], wmax[1][rq]), fmaxf(wmax[2][rq], wmax[3][rq]));
    const int e = block_exp(amax);
    u32x4 out[2];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const f4v x = {T[rq][seg * 32 + 4 * j], T[rq][seg * 32 + 4 * j + 1], T[rq][seg * 32 + 4 * j + 2],
                       T[rq][seg * 32 + 4 * j + 3]};
        out[j / 4][j % 4] = pack4(x, e);
    }
    u32x4 *dst = reinterpret_cast<u32x4 *>(Q + r * (KB * kQB) + k0 + seg * 32);
    dst[0] = out[0];
    dst[1] = out[1];
    if (seg == 0) S[scale_byte(r, kb, Rp)] = (uint8_t)(127 + e);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
fp8_gemm_kernel(const uint8_t *__restrict__ Qa, const uint8_t *__restrict__ Qb, const uint32_t *__restrict__ Sa,
                const uint32_t *__restrict__ Sb, int64_t Kp, int64_t Mp, int64_t Np, int tiles_m, int tiles_n,
                float *__restrict__ C, int64_t ldc, int64_t M, int64_t N, float alpha, float beta)
{
    __shared__ __attribute__((aligned(16))) uint8_t lds[2][2][kTile * kRow];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lg = lane >> 4;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;

    // XCD-aware grouped tile order, as in the FP32/FP64 tile kernels.
    int pid = blockIdx.x;
    const int nwg = gridDim.x;
    if ((nwg & 7) == 0) pid = (pid & 7) * (nwg >> 3) + (pid >> 3);
    const int width = 8 * tiles_n;
    const int first_m = (pid / width) * 8;
    const int gsz = min(tiles_m - first_m, 8);
    const int tm = first_m + (pid % width) % gsz, tn = (pid % width) / gsz;
    const int64_t m0 = (int64_t)tm * kTile, n0 = (int64_t)tn * kTile;

    // staging: 128 rows x 128 bytes per operand = 1024 16-byte vectors, 4 per thread
    u32x4 ra[4], rb[4];
    auto gload = [&](int64_t k0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int c = tid + i * 256, row = c >> 3, col = (c & 7) * 16;
            ra[i] = *reinterpret_cast<const u32x4 *>(Qa + (m0 + row) * Kp + k0 + col);
            rb[i] = *reinterpret_cast<const u32x4 *>(Qb + (n0 + row) * Kp + k0 + col);
        }
    };
    auto lstore = [&](int st) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int c = tid + i * 256, row = c >> 3, col = (c & 7) * 16;
            *reinterpret_cast<u32x4 *>(&lds[st][0][row * kRow + col]) = ra[i];
            *reinterpret_cast<u32x4 *>(&lds[st][1][row * kRow + col]) = rb[i];
        }
    };
    // scale words of this lane's four A rows and four B columns in k block kb
    const int64_t a_grp = (m0 + wm) / kRowBlock, b_grp = (n0 + wn) / kRowBlock;
    const int64_t a_groups = Mp / kRowBlock, b_groups = Np / kRowBlock;
    auto sload = [&](int64_t kb, uint32_t &sa, uint32_t &sb) {
        sa = Sa[(kb * a_groups + a_grp) * 16 + lr];
        sb = Sb[(kb * b_groups + b_grp) * 16 + lr];
    };

    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4{0, 0, 0, 0};

    const int nk = (int)(Kp / kQB);
    uint32_t sa, sb, sa_next = 0, sb_next = 0;
    gload(0);
    sload(0, sa, sb);
    lstore(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) {
            gload((int64_t)(kt + 1) * kQB);
            sload(kt + 1, sa_next, sb_next);
        }
        i8v fa[4], fb[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint8_t *pa = &lds[cur][0][(wm + i * 16 + lr) * kRow + lg * 32];
            const uint8_t *pb = &lds[cur][1][(wn + i * 16 + lr) * kRow + lg * 32];
            const u32x4 a0 = *reinterpret_cast<const u32x4 *>(pa), a1 = *reinterpret_cast<const u32x4 *>(pa + 16);
            const u32x4 b0 = *reinterpret_cast<const u32x4 *>(pb), b1 = *reinterpret_cast<const u32x4 *>(pb + 16);
            fa[i] = i8v{(int)a0[0], (int)a0[1], (int)a0[2], (int)a0[3], (int)a1[0], (int)a1[1], (int)a1[2], (int)a1[3]};
            fb[i] = i8v{(int)b0[0], (int)b0[1], (int)b0[2], (int)b0[3], (int)b1[0], (int)b1[1], (int)b1[2], (int)b1[3]};
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int sai = (int)((sa >> (8 * i)) & 0xFFu);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int sbj = (int)((sb >> (8 * j)) & 0xFFu);
                acc[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(fa[i], fb[j], acc[i][j], 0, 0, 0, sai,
                                                                             0, sbj);
            }
        }
        if (more) {
            lstore(cur ^ 1);
            sa = sa_next;
            sb = sb_next;
        }
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = m0 + wm + i * 16 + 4 * lg + r, col = n0 + wn + j * 16 + lr;
                if (row < M && col < N) {
                    float v = alpha * acc[i][j][r];
                    if (beta != 0.0f) v += beta * C[row * ldc + col];
                    C[row * ldc + col] = v;
                }
            }
}

// The direct-to-LDS GEMM, TM x TN tiles of NW waves in a 2 x (NW / 2) grid
// ((TM / 2) x (2 TN / NW) each).  Per K step of 128 a workgroup needs TM + TN
// rows x 128 bytes; they go global -> LDS by direct-to-LDS buffer loads (lane
// l's 16 bytes land at the block base + 16 l), 1 KiB each, into a ring of NS
// LDS stages, NS - 1 steps ahead.  A 1 KiB block holds 8 rows; the 16-byte
// chunk c of row r sits in slot 8 r + (c ^ x(r)), x(r) = 4 r_3 + 2 r_1 + r_2
// (r_i: bit i of r), so lane l of a load of block b reads row l / 8, chunk
// (l % 8) ^ (4 (b % 2) + 2 ((l / 16) % 2) + l / 32).  A fragment read (lane
// l: row l % 16, chunk 2 (l / 16) or 2 (l / 16) + 1) then puts every group
// of 8, 16 or 32 consecutive lanes on distinct four-byte banks.  Measured
// on MI355X with a probe of these reads alone: 125 TB/s against 73 TB/s for
// the swizzles c ^ (r & 7) and c ^ ((r >> 1) & 7) (each of which left
// SQ_LDS_BANK_CONFLICT at 4 cycles per read) and 39 TB/s with none.
// Each wave's scale words of a step (its rows' and its columns' 64-row
// groups; one word holds the scales of four 16-row fragments) go to the
// wave's own 512 bytes of the stage by two more direct-to-LDS loads, so no
// register waits on a load across steps.
template <int TM, int TN, int NW, int NS>
struct BigTile {
    static constexpr int BLOCKS_A = TM / 8, BLOCKS = (TM + TN) / 8;  // 1 KiB blocks per stage
    static constexpr int LOADS = BLOCKS / NW;                        // per wave
    static constexpr int WC = NW / 2;                                // wave columns
    static constexpr int MI = TM / 32, NJ = TN / (16 * WC);          // MFMA tiles per wave
    static constexpr int GA = (MI + 3) / 4, GB = (NJ + 3) / 4;       // scale words per lane and operand
    static constexpr int VM = LOADS + 2;                             // vector-memory ops per stage
    static constexpr int STAGE = BLOCKS * 1024 + NW * 512;           // bytes per stage
    static constexpr int WPE = 2 * NS * STAGE <= 160 * 1024 ? 2 : 1;  // workgroups per CU the LDS allows
    static_assert(BLOCKS % NW == 0 && NJ >= 1 && (TM / 2) % 64 == 0 && (TN / WC) % 64 == 0, "tile shape");
    static_assert(NS >= 2 && NS * STAGE <= 160 * 1024 && (NS - 2) * VM <= 63, "stages");
};

// s_waitcnt vmcnt(N) through the builtin, so the compiler's own wait
// insertion sees it (gfx9 encoding: vmcnt bits 3:0 and 15:14; expcnt and
// lgkmcnt left at their maxima).
template <int N>
__device__ __forceinline__ void wait_vmcnt()
{
    static_assert(N >= 0 && N <= 63, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// f(integral_constant<int, 0>()), ..., f(integral_constant<int, N - 1>()).
template <typename F, int... I>
__device__ __forceinline__ void static_for_seq(F &&f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>()), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    static_for_seq(f, std::make_integer_sequence<int, N>());
}

// This file is compiled with the MFMAs in their VGPR form (see the
// Makefile): otherwise, with the register budget of one wave per SIMD, the
// compiler parks the accumulators in AGPRs and copies them around every MFMA.
template <int TM, int TN, int NW, int NS>
__global__ void __launch_bounds__(64 * NW)
__attribute__((amdgpu_waves_per_eu(BigTile<TM, TN, NW, NS>::WPE * NW / 4, BigTile<TM, TN, NW, NS>::WPE * NW / 4)))
fp8_big_kernel(const uint8_t *__restrict__ Qa, const uint8_t *__restrict__ Qb, const uint32_t *__restrict__ Sa,
               const uint32_t *__restrict__ Sb, int64_t Kp, int64_t Mp, int64_t Np, int tiles_m, int tiles_n,
               float *__restrict__ C, int64_t ldc, int64_t M, int64_t N, float alpha, float beta)
{
    using Bt = BigTile<TM, TN, NW, NS>;
    constexpr int MI = Bt::MI, NJ = Bt::NJ, GA = Bt::GA, GB = Bt::GB;
    constexpr int kStage = Bt::STAGE;
    __shared__ __attribute__((aligned(1024))) uint8_t lds[NS * kStage];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lr = lane & 15, lg = lane >> 4;
    const int wm = (wave / Bt::WC) * (TM / 2), wn = (wave % Bt::WC) * (TN / Bt::WC);

    // XCD-aware grouped tile order, as in the other GEMM kernels.
    int pid = blockIdx.x;
    const int nwg = gridDim.x;
    if ((nwg & 7) == 0) pid = (pid & 7) * (nwg >> 3) + (pid >> 3);
    const int width = 8 * tiles_n;
    const int first_m = (pid / width) * 8;
    const int gsz = min(tiles_m - first_m, 8);
    const int tm = first_m + (pid % width) % gsz, tn = (pid % width) / gsz;
    const int64_t m0 = (int64_t)tm * TM, n0 = (int64_t)tn * TN;

    const int nk = (int)(Kp / kQB);
    // lane l of a load of an even (odd) block: row l / 8 of the block, chunk
    // (l % 8) ^ x with x = 2 ((l / 16) % 2) + l / 32 (plus 4)
    static_assert(Bt::LOADS % 2 == 0 && Bt::BLOCKS_A % 2 == 0, "block parity is the load index's");
    const int xl = (((lane >> 4) & 1) << 1) | (lane >> 5);
    const uint32_t voff[2] = {(uint32_t)((lane >> 3) * Kp + (((lane & 7) ^ xl) * 16)),
                              (uint32_t)((lane >> 3) * Kp + (((lane & 7) ^ (4 | xl)) * 16))};
    const int64_t a_groups = Mp / kRowBlock, b_groups = Np / kRowBlock;
    const int64_t a_grp = (m0 + wm) / kRowBlock, b_grp = (n0 + wn) / kRowBlock;
    // lane l of a scale load: word l of the wave's GA (GB) consecutive
    // 16-word groups; the lanes past them repeat the last word
    const uint32_t soff_a = (uint32_t)(min(lane, 16 * GA - 1) * 4), soff_b = (uint32_t)(min(lane, 16 * GB - 1) * 4);
    const int scale_at = Bt::BLOCKS * 1024 + wave * 512;  // this wave's scale words in a stage
    // Loads of K step kb into stage S: this wave's share of the blocks, then
    // its scale words (Bt::VM vector-memory operations).
    auto issue = [&](int kb, auto s_c) {
        constexpr int S = decltype(s_c)::value;
        const int64_t k = min(kb, nk - 1);
        uint8_t *dst = lds + S * kStage;
#pragma unroll
        for (int q = 0; q < Bt::LOADS; ++q) {
            const int b = wave * Bt::LOADS + q;  // wave-uniform
            const bool isb = b >= Bt::BLOCKS_A;
            const int rows = (isb ? b - Bt::BLOCKS_A : b) * 8;
            const uint8_t *base = (isb ? Qb + (n0 + rows) * Kp : Qa + (m0 + rows) * Kp) + k * kQB;
            const __amdgpu_buffer_rsrc_t rs =
                __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t *>(base), (short)0, 0x7fffffff, 0x00020000);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void *)(dst + b * 1024),
                                                     16, voff[q & 1], 0, 0, 0);
        }
        const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<uint32_t *>(Sa + (k * a_groups + a_grp) * 16), (short)0, 0x7fffffff, 0x00020000);
        __builtin_amdgcn_raw_ptr_buffer_load_lds(ra, (__attribute__((address_space(3))) void *)(dst + scale_at), 4,
                                                 soff_a, 0, 0, 0);
        const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<uint32_t *>(Sb + (k * b_groups + b_grp) * 16), (short)0, 0x7fffffff, 0x00020000);
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rb, (__attribute__((address_space(3))) void *)(dst + scale_at + 256),
                                                 4, soff_b, 0, 0, 0);
    };

    f4 acc[MI][NJ];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[i][j] = f4{0, 0, 0, 0};

    // Step kt on stage S = kt % NS: wait for this wave's loads of stage kt
    // (the NS - 2 later stages' may stay in flight), barrier (every wave's
    // have landed; every wave is done with the stage step kt - 1 read), load
    // step kt + NS - 1 into that stage, read this wave's fragments, issue its
    // MFMAs.  Past the last step a step reloads the last one (the same count
    // of loads every step); those stages are never read.
    auto step = [&](int kt, auto s_c) {
        constexpr int S = decltype(s_c)::value;
        __builtin_amdgcn_sched_barrier(0);
        wait_vmcnt<(NS - 2) * Bt::VM>();
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
        issue(kt + NS - 1, std::integral_constant<int, (S + NS - 1) % NS>());
        // 8 waves (two per SIMD, one workgroup): the wave in its fragment
        // reads and MFMAs takes issue priority over the other's loads and
        // barrier wait.  At 8192^3 the whole form measured 0.682-0.683 ms
        // against 0.689-0.708 (bit-identical); for the 128-tile kernel, two
        // workgroups per CU, it was mixed (2048 x 2048 x 8192 2-5 % slower).
        if constexpr (NW == 8) __builtin_amdgcn_s_setprio(1);
        const uint8_t *base = lds + S * kStage;
        auto frag = [&](const uint8_t *op, int row) -> i8v {
            const int c0 = 2 * lg, sw = (((row >> 3) & 1) << 2) | (((row >> 1) & 1) << 1) | ((row >> 2) & 1);
            const u32x4 x0 = *reinterpret_cast<const u32x4 *>(op + row * 128 + ((c0 ^ sw) * 16));
            const u32x4 x1 = *reinterpret_cast<const u32x4 *>(op + row * 128 + (((c0 + 1) ^ sw) * 16));
            return i8v{(int)x0[0], (int)x0[1], (int)x0[2], (int)x0[3], (int)x1[0], (int)x1[1], (int)x1[2], (int)x1[3]};
        };
        // op(B)'s fragments and the scale words stay in registers for the
        // step; op(A)'s fragments are read kAhead row tiles ahead of the NJ
        // MFMAs that use them, in the order the group barriers pin: left to
        // itself the compiler kept two of them and waited on each read right
        // after issuing it.
        const uint32_t *sw = reinterpret_cast<const uint32_t *>(base + scale_at);
        uint32_t sc[GA + GB];
#pragma unroll
        for (int g = 0; g < GA; ++g) sc[g] = sw[16 * g + lr];
#pragma unroll
        for (int g = 0; g < GB; ++g) sc[GA + g] = sw[64 + 16 * g + lr];
        i8v fb[NJ], fa[MI];
#pragma unroll
        for (int j = 0; j < NJ; ++j) fb[j] = frag(base + TM * 128, wn + 16 * j + lr);
        constexpr int kAhead = MI < 2 ? MI : 2;
#pragma unroll
        for (int i = 0; i < kAhead; ++i) fa[i] = frag(base, wm + 16 * i + lr);
#pragma unroll
        for (int i = 0; i < MI; ++i) {
            if (i + kAhead < MI) fa[i + kAhead] = frag(base, wm + 16 * (i + kAhead) + lr);
            const int sai = (int)((sc[i >> 2] >> (8 * (i & 3))) & 0xFFu);
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                const int sbj = (int)((sc[GA + (j >> 2)] >> (8 * (j & 3))) & 0xFFu);
                acc[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(fa[i], fb[j], acc[i][j], 0, 0, 0, sai, 0,
                                                                             sbj);
            }
        }
        // DS reads: the scale words, op(B)'s fragments and the first op(A)
        // fragments (two 16-byte reads each); then per row tile the reads of
        // the fragment kAhead ahead and the row tile's MFMAs.
        __builtin_amdgcn_sched_group_barrier(0x100, 2 + 2 * NJ + 2 * kAhead, 0);
#pragma unroll
        for (int i = 0; i < MI; ++i) {
            if (i + kAhead < MI) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, NJ, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (NW == 8) __builtin_amdgcn_s_setprio(0);
    };
    static_for<NS - 1>([&](auto s_c) { issue(decltype(s_c)::value, s_c); });
    // NS steps per iteration, so every stage index is a constant
    int kt = 0;
    for (; kt + NS <= nk; kt += NS) static_for<NS>([&](auto s_c) { step(kt + decltype(s_c)::value, s_c); });
    static_for<NS>([&](auto s_c) {
        if (kt + decltype(s_c)::value < nk) step(kt + decltype(s_c)::value, s_c);
    });
    wait_vmcnt<0>();  // no load into this workgroup's LDS may be in flight when it exits

#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = m0 + wm + i * 16 + 4 * lg + r, col = n0 + wn + j * 16 + lr;
                if (row < M && col < N) {
                    float v = alpha * acc[i][j][r];
                    if (beta != 0.0f) v += beta * C[row * ldc + col];
                    C[row * ldc + col] = v;
                }
            }
}

// Tiles of the direct-to-LDS FP8 GEMM: 128 x 128 with 4 waves (two
// workgroups per CU) in general, 256 x 256 with 8 waves (one per CU) when a
// launch has at least four of those per CU.  Measured on MI355X, all
// bit-identical: at 8192^3 the whole form took 0.69-0.71 ms with 256 tiles
// against 0.72-0.76 with 128 tiles; at 4096^3 the two were equal, at
// 2048 x 2048 x 8192 and 3000 x 8192 x 2048 the 256 tiles 10-50 % slower.
// A ring of 3 or 4 stages at one workgroup per CU ran 0.92-0.93 ms at 8192^3:
// the second workgroup keeps the matrix cores busy across each barrier.
constexpr int kBigS = 2;
constexpr int kWideTile = 256;

int fp8_tile(int64_t M, int64_t N)
{
    const int64_t wide = ((M + kWideTile - 1) / kWideTile) * ((N + kWideTile - 1) / kWideTile);
    return wide >= 4 * (int64_t)num_cus() ? kWideTile : 128;
}

struct Fp8Layout {
    int64_t Mp, Np, Kp, KB;
    uint8_t *Qa, *Qb, *Sa, *Sb;
    size_t bytes;
};

size_t up256(size_t b) { return (b + 255) / 256 * 256; }

// Quantized operands padded to whole tiles of edge `tile` (a multiple of
// the 128 of the register-staged kernel).
Fp8Layout fp8_layout(int64_t M, int64_t N, int64_t K, int tile, void *base)
{
    Fp8Layout L;
    L.Mp = (M + tile - 1) / tile * tile;
    L.Np = (N + tile - 1) / tile * tile;
    L.KB = (K + kQB - 1) / kQB;
    L.Kp = L.KB * kQB;
    uint8_t *p = (uint8_t *)base;
    size_t off = 0;
    L.Qa = p + off;
    off += up256((size_t)(L.Mp * L.Kp));
    L.Qb = p + off;
    off += up256((size_t)(L.Np * L.Kp));
    L.Sa = p + off;
    off += up256((size_t)(L.Mp * L.KB));
    L.Sb = p + off;
    off += up256((size_t)(L.Np * L.KB));
    L.bytes = off;
    return L;
}

hipError_t quantize(const float *X, int64_t ld, bool k_contig, int64_t R, int64_t K, int64_t Rp, int64_t KB,
                    uint8_t *Q, uint8_t *S, hipStream_t s)
{
    const bool vec = ((uintptr_t)X % 16) == 0 && ld % 4 == 0;
    if (KB > 65535 || Rp / 64 > 0x7fffffffLL) return hipErrorInvalidValue;  // grid limits
    if (k_contig) {
        const int64_t per_wave = 2 * kQuantBlocks, waves = (Rp * KB + per_wave - 1) / per_wave;
        const dim3 grid((unsigned)((waves + 3) / 4));
        if (vec) hipLaunchKernelGGL(quant_kc_kernel<true>, grid, dim3(256), 0, s, X, ld, R, K, Rp, KB, Q, S);
        else hipLaunchKernelGGL(quant_kc_kernel<false>, grid, dim3(256), 0, s, X, ld, R, K, Rp, KB, Q, S);
    } else {
        const dim3 grid((unsigned)(Rp / 64), (unsigned)KB);
        if (vec) hipLaunchKernelGGL(quant_mc_kernel<true>, grid, dim3(256), 0, s, X, ld, R, K, Rp, KB, Q, S);
        else hipLaunchKernelGGL(quant_mc_kernel<false>, grid, dim3(256), 0, s, X, ld, R, K, Rp, KB, Q, S);
    }
    return hipGetLastError();
}

}  // namespace

size_t fp8_workspace_bytes(int64_t M, int64_t N, int64_t K)
{
    return fp8_layout(M, N, K, fp8_tile(M, N), nullptr).bytes;
}

hipError_t launch_fp8_gemm(const Gemm<float> &g, void *work, hipStream_t s)
{
    const int tile = fp8_tile(g.M, g.N);
    const Fp8Layout L = fp8_layout(g.M, g.N, g.K, tile, work);
    // op(A) rows are K-contiguous unless A is transposed; op(B) columns are
    // K-contiguous exactly when B is transposed (stored N x K).
    hipError_t e = quantize(g.A, g.lda, !g.ta, g.M, g.K, L.Mp, L.KB, L.Qa, L.Sa, s);
    if (e == hipSuccess) e = quantize(g.B, g.ldb, g.tb, g.N, g.K, L.Np, L.KB, L.Qb, L.Sb, s);
    if (e != hipSuccess) return e;
    // the big kernel's lane offsets span 8 rows of the quantized operands
    if (8.0 * (double)L.Kp + 128.0 < 2147483647.0) {
        if ((L.Mp / tile) * (L.Np / tile) > 0x7fffffffLL) return hipErrorInvalidValue;
        const int tiles_m = (int)(L.Mp / tile), tiles_n = (int)(L.Np / tile);
        const dim3 grid((unsigned)(tiles_m * tiles_n));
        if (tile == kWideTile)
            hipLaunchKernelGGL((fp8_big_kernel<kWideTile, kWideTile, 8, kBigS>), grid, dim3(512), 0, s, L.Qa, L.Qb,
                               (const uint32_t *)L.Sa, (const uint32_t *)L.Sb, L.Kp, L.Mp, L.Np, tiles_m, tiles_n, g.C,
                               g.ldc, g.M, g.N, g.alpha, g.beta);
        else
            hipLaunchKernelGGL((fp8_big_kernel<128, 128, 4, kBigS>), grid, dim3(256), 0, s, L.Qa, L.Qb,
                               (const uint32_t *)L.Sa, (const uint32_t *)L.Sb, L.Kp, L.Mp, L.Np, tiles_m, tiles_n, g.C,
                               g.ldc, g.M, g.N, g.alpha, g.beta);
        return hipGetLastError();
    }
    if ((L.Mp / kTile) * (L.Np / kTile) > 0x7fffffffLL) return hipErrorInvalidValue;
    const int tiles_m = (int)(L.Mp / kTile), tiles_n = (int)(L.Np / kTile);
    hipLaunchKernelGGL(fp8_gemm_kernel, dim3((unsigned)(tiles_m * tiles_n)), dim3(256), 0, s, L.Qa, L.Qb,
                       (const uint32_t *)L.Sa, (const uint32_t *)L.Sb, L.Kp, L.Mp, L.Np, tiles_m, tiles_n, g.C,
                       g.ldc, g.M, g.N, g.alpha, g.beta);
    return hipGetLastError();
}

}  // namespace blasdsp
