// gemm_skinny.hip — bandwidth-bound GEMM for the irregular shapes: shortA
// (M small, op(B) K x N streamed) and thinB (N small, op(A) M x K streamed).
//
// What it replaces.  irr_shortA_{d,s}gemm_* keep the small A whole on chip and
// split N across DSP cores; irr_thinB_{d,s}gemm_* keep B whole, split M, and
// compute C^T through "transpose + shortA + transpose"
// (irr_shortA_dgemm_release_V0.0/device_code/irr_shortA_dgemm_kernel.c:1607-2677,
// irr_thinB_sgemm_release_V1.2/device_code/irr_thinB_sgemm_kernel.c:2910-5157).
//
// One kernel serves both.  It computes O = S * L where S ("small", Ms x K,
// Ms <= 64) is the small operand and L ("large", K x Nl) is the streamed one:
//   shortA:  S = op(A), L = op(B), O = C           (stored C[ms][nl])
//   thinB:   S = op(B)^T, L = op(A)^T, O = C^T     (stored C[nl][ms])
// so thinB is shortA with the operand roles swapped and a transposed store —
// the same identity the reference uses, with the transposes absorbed into the
// index maps instead of executed.
//
// Each wave owns a 64-column panel of O (all Ms rows, MT = ceil(Ms/16) = 1,
// 2, 3 or 4 MFMA row tiles x 4 column tiles) and walks K in blocks of one 128-byte line per row
// (32 FP32 or 16 FP64 values; two lines for K-contiguous FP64 L).  Every
// element of L is loaded exactly once, with non-temporal 16-byte loads.  The
// MFMA K index is permuted so that each lane's vector holds whole operands:
//   * L stored k-major ([k][nl], nl contiguous): lane (c, g) = (l%16, l/16)
//     loads row k0 + g*R + j, columns n0 + V*c .. + V-1 at virtual step j
//     (R = block / 4) straight into operand registers; element e of the
//     vector is the B operand of column tile e, which therefore covers
//     columns n0 + V*c + e (an interleaved column map undone at the store);
//   * L stored nl-major ([nl][k], k contiguous): each load instruction reads
//     whole row pieces of the block, which the wave writes to its own LDS
//     rows and reads back as operands (row n0+16t+c, V consecutive k per
//     vector, the four lane groups side by side).
// S is read with the same k pairing as L (k_of below; a run along k when S
// is k-contiguous); it is tiny (<= 64 x K) and served from L1/L2.
//
// Full blocks are buffer loads from wave-uniform bases with no lane
// branches, kept in flight in a ring of register blocks (load_full and the
// ring in skinny_kernel); ragged blocks take bounds-checked loads.
//
// When the Nl panels alone leave the GPU's workgroup slots poorly filled, K
// is split into slices (grid.y, skinny_pick_splitk); each slice writes its
// partial O to a workspace and a second kernel sums the slices in fixed order
// (deterministic, no atomics) and applies C = alpha * sum + beta * C.
#include <type_traits>

#include "internal.h"

namespace blasdsp {
namespace {

template <typename T>
struct SK;

template <>
struct SK<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    typedef double vec_t __attribute__((ext_vector_type(2)));
    static constexpr int V = 2;
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // f64 16x16 output: lane group g, register r hold row g + 4 r.
    static __device__ __forceinline__ int out_row(int g, int r) { return g + 4 * r; }
};

template <>
struct SK<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    typedef float vec_t __attribute__((ext_vector_type(4)));
    static constexpr int V = 4;
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // f32 16x16 output: lane group g, register r hold row 4 g + r.
    static __device__ __forceinline__ int out_row(int g, int r) { return 4 * g + r; }
};

template <typename T>
struct SkinnyArgs {
    const T *S;
    int64_t lds;
    const T *L;
    int64_t ldl;
    T *C;
    int64_t ldc;
    T *W;             // split partials [split][Ms][Nl], used when splitk > 1
    int64_t Ms, Nl, K;
    int64_t kslice;   // K per slice, a multiple of the K block
    int splitk;
    T alpha, beta;
    bool l_vec, s_vec;
    bool c_vec;       // C 16-byte aligned, ldc a multiple of the vector
};

constexpr int kWaves = 4;      // panels per workgroup, 64 columns of O each
constexpr int kPanel = 64;     // O columns per wave

// Waves that share one panel inside a workgroup, each walking its own part
// of the K range (1 or 2).  With 2, a workgroup is 8 waves, two per SIMD on
// the same four panels: while one wave of a SIMD waits for its loads the
// other issues MFMAs, and the two partial sums meet in LDS at the end, so
// the streamed operand is still read exactly once.  Measured on MI355X at
// K = 8192, long side 65536: 32 rows of S (where the matrix cores must run
// at ~65 % of peak to keep pace with the stream) gain 1-16 % per case, except
// FP64 with k-major L, which is 1-3 % slower and keeps one wave.  At 8 and 16
// rows only FP32 with K-contiguous L gains (0.2-3.5 % per case, interleaved
// A/B); the other 8/16-row cases gain nothing and keep one wave.
template <typename T>
__host__ __device__ constexpr int skinny_ks(int mt, bool l_kc)
{
    return (mt == 2 && (sizeof(T) == 4 || l_kc)) || (mt == 1 && sizeof(T) == 4 && l_kc) ? 2 : 1;
}

// 128-byte lines of each K-contiguous row of L per block.  A wave's block
// covers 64 rows, and with one line per row the concurrent waves touch 128
// bytes of many distinct rows at a time.  Two lines measured 1-3 % faster for
// FP64 (16 and 32 rows of S, one wave per SIMD) and 0.5-1.4 % slower for FP32
// at K = 8192; 64-row S keeps one line (the larger block does not fit its
// registers), and so do K-split waves (two waves share a SIMD's registers).
constexpr int kLinesKC = 2;
template <typename T>
__host__ __device__ constexpr int lines_of(bool l_kc, int mt)
{
    return l_kc && mt <= 2 && sizeof(T) == 8 && skinny_ks<T>(mt, l_kc) == 1 ? kLinesKC : 1;
}

// K per block: LINES 128-byte lines (32 FP32 / 16 FP64 per line).
template <typename T, int LINES>
struct Blk {
    static constexpr int V = SK<T>::V;            // elements per 16-byte vector
    static constexpr int KB = LINES * 128 / (int)sizeof(T);
    static constexpr int RUN = KB / 4;            // k-run per lane group
    static constexpr int NSTEP = RUN;             // MFMA k-steps per block
    static constexpr int RV = RUN / V;            // vectors per run
    static constexpr int NLV = 4 * RV;            // L vectors per lane per block
    static constexpr int CPR = KB / V;            // 16-byte chunks per K-contiguous row
    static constexpr int RPI = 64 / CPR;          // rows per load instruction (K-contiguous)
};

// Column of O held by lane column c of MFMA column tile t.
template <typename T, bool L_KC>
__device__ __forceinline__ int col_of(int t, int c)
{
    constexpr int V = SK<T>::V;
    if (L_KC) return 16 * t + c;
    // t = h * V + e: vector half h, element e
    return (t / V) * 16 * V + V * c + (t % V);
}

// k paired with lane group g at virtual MFMA step j of the block at k0, the
// same on both operands.  K-contiguous L: vector q of every lane group sits
// side by side (k0 + 4 V q + V g + e), so one load instruction reads 64
// contiguous bytes of each of its 16 rows.  k-major L: lane group g walks its
// own run (k0 + RUN g + j), one row of L per step.
template <typename T, bool L_KC>
__device__ __forceinline__ int64_t k_of(int64_t k0, int g, int j)
{
    constexpr int V = Blk<T, 1>::V, RUN = Blk<T, 1>::RUN;
    if (L_KC) return k0 + (j / V) * 4 * V + g * V + (j % V);
    return k0 + g * RUN + j;
}

// Registers for one K block of one wave; step j pairs lane group g with
// k_of(k0, g, j) on both operands.
template <typename T, int MT, bool S_KC, bool L_KC>
struct Block {
    using vec_t = typename SK<T>::vec_t;
    using B_ = Blk<T, lines_of<T>(L_KC, MT)>;
    static constexpr int V = B_::V;
    vec_t l[B_::NLV];
    // S: k-contiguous -> runs as vectors; m-contiguous -> scalars per step
    vec_t s_run[MT][S_KC ? B_::RV : 1];
    T s_scl[MT][S_KC ? 1 : B_::NSTEP];

    // L operand of column tile t at virtual step j.
    __device__ __forceinline__ T lop(int t, int j) const
    {
        if (L_KC) return l[t * B_::RV + j / V][j % V];  // t's run, vector j / V
        return l[j * (4 / V) + t / V][t % V];           // step j's row, half t / V
    }
    __device__ __forceinline__ T sop(int mt, int j) const
    {
        if (S_KC) return s_run[mt][j / V][j % V];
        return s_scl[mt][j];
    }
};

// The streamed operand is read exactly once: non-temporal loads keep it from
// displacing the small operand in the caches (a plain 4 GiB read streams at
// 6.1-6.3 TB/s on MI355X, the same read non-temporal at 6.96-6.98 TB/s).
// Full blocks read whole row pieces and always use them (kAuxNT below).  In
// ragged blocks they are used where each instruction covers whole 64-byte
// runs of a row (L stored with the long side contiguous); in the operand
// layout of K-contiguous L an instruction covers half of each of 16 rows'
// 128-byte lines, where non-temporal loads measured 10-15 % slower than
// plain ones, so those stay plain.
template <typename V>
__device__ __forceinline__ V stream_load(const V *p)
{
    return __builtin_nontemporal_load(p);
}

// Byte offsets of one lane from the wave-uniform bases of the unmasked loads.
// They do not change from block to block: the loads are buffer loads whose
// resource (the base) is rebuilt in scalar registers per block, so each
// operand needs one offset register in total (the host checks that the
// offsets fit: skinny_fits).
template <int MT>
struct LaneOff {
    uint32_t l;        // streamed operand
    uint32_t s[MT];    // small operand, per 16-row tile
};

// k_of(k0, g, j) = k0 + kgrp(g) + kstep(j): the lane-group part and the step
// part of the permuted k index.
template <typename T, bool L_KC>
__device__ __forceinline__ int kgrp(int g)
{
    return L_KC ? g * Blk<T, 1>::V : g * Blk<T, 1>::RUN;
}
template <typename T, bool L_KC>
__host__ __device__ constexpr int kstep(int j)
{
    return L_KC ? (j / Blk<T, 1>::V) * 4 * Blk<T, 1>::V + j % Blk<T, 1>::V : j;
}

template <typename T, int MT, int SM, bool L_KC>
__device__ __forceinline__ LaneOff<MT> lane_offsets(const SkinnyArgs<T> &a, int c, int g)
{
    using B_ = Blk<T, lines_of<T>(L_KC, MT)>;
    constexpr int V = B_::V, RUN = B_::RUN;
    constexpr int64_t s = sizeof(T);
    const int lane = g * 16 + c;
    LaneOff<MT> o;
    if (L_KC) o.l = (uint32_t)(((int64_t)(lane / B_::CPR) * a.ldl + (lane % B_::CPR) * V) * s);
    else o.l = (uint32_t)(((int64_t)g * RUN * a.ldl + V * c) * s);
    const int64_t kg = kgrp<T, L_KC>(g);
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        // Rows past Ms read row Ms - 1.  An MFMA output row depends only on
        // the same row of the A operand, and output rows past Ms are never
        // stored, so those rows need no zeroing.
        const int64_t ms = mt * 16 + c;
        const int64_t mr = ms < a.Ms ? ms : a.Ms - 1;
        o.s[mt] = (uint32_t)((SM != 0 ? mr * a.lds + kg : kg * a.lds + mr) * s);
    }
    return o;
}

// Buffer loads from a wave-uniform base: raw buffer, no stride, records up to
// 2^31 bytes (lane offsets stay below that), the CDNA data-format word.
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const T *base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(base), (short)0, 0x7fffffff, 0x00020000);
}
constexpr int kAuxNT = 2;  // cache policy: non-temporal
template <typename T, int AUX>
__device__ __forceinline__ typename SK<T>::vec_t bload_vec(const T *base, uint32_t off)
{
    return __builtin_bit_cast(typename SK<T>::vec_t, __builtin_amdgcn_raw_buffer_load_b128(rsrc(base), off, 0, AUX));
}
template <typename T>
__device__ __forceinline__ T bload(const T *base, uint32_t off)
{
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rsrc(base), off, 0, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rsrc(base), off, 0, 0));
}

// One full K block (k0 .. k0 + KB of a panel inside Nl) with no bounds
// checks and no lane branches: every block issues the same loads, so the
// wait before a block's MFMAs counts exactly the loads issued after it.
// n0 and k0 are wave-uniform.
// WITH_S false: L only (S comes from the workgroup's shared chunk).
template <typename T, int MT, int SM, bool L_KC, bool WITH_S = true>
__device__ __forceinline__ void load_full(Block<T, MT, SM != 0, L_KC> &b, const SkinnyArgs<T> &a,
                                          const LaneOff<MT> &o, int64_t n0, int64_t k0)
{
    using B_ = Blk<T, lines_of<T>(L_KC, MT)>;
    constexpr int V = B_::V;
    if (L_KC) {
        // vector i: 16 bytes of row n0 + RPI i + lane / CPR, chunk lane % CPR;
        // each instruction reads whole KB-long row pieces of RPI rows, which
        // stage_block() turns into MFMA operands through LDS
#pragma unroll
        for (int i = 0; i < B_::NLV; ++i)
            b.l[i] = bload_vec<T, kAuxNT>(a.L + (n0 + B_::RPI * i) * a.ldl + k0, o.l);
    } else {
        // step j: row k_of(k0, g, j), columns n0 + h*16V + V*c .. + V-1
#pragma unroll
        for (int j = 0; j < B_::NSTEP; ++j)
#pragma unroll
            for (int h = 0; h < 4 / V; ++h)
                b.l[j * (4 / V) + h] = bload_vec<T, kAuxNT>(a.L + (k0 + j) * a.ldl + n0 + h * 16 * V, o.l);
    }
    if constexpr (!WITH_S) return;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        if (SM != 0) {
#pragma unroll
            for (int q = 0; q < B_::RV; ++q) {
                const T *base = a.S + k0 + kstep<T, L_KC>(q * V);
                if constexpr (SM == 2) {
                    b.s_run[mt][q] = bload_vec<T, 0>(base, o.s[mt]);
                } else {
#pragma unroll
                    for (int e = 0; e < V; ++e) b.s_run[mt][q][e] = bload<T>(base + e, o.s[mt]);
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < B_::NSTEP; ++j)
                b.s_scl[mt][j] = bload<T>(a.S + (k0 + kstep<T, L_KC>(j)) * a.lds, o.s[mt]);
        }
    }
}

// A ragged block (partial panel or partial K block), every element
// bounds-checked, L read straight into the operand layout.  S access mode SM:
// 0 = S stored m-contiguous (one scalar per step), 1 = K-contiguous read as
// scalars (S not 16-byte aligned, or lds not a multiple of the vector), 2 =
// K-contiguous read as 16-byte vectors.
template <typename T, int MT, int SM, bool L_KC>
__device__ __forceinline__ void load_tail(Block<T, MT, SM != 0, L_KC> &b, const SkinnyArgs<T> &a,
                                          int64_t n0, int64_t k0, int64_t kend, int c, int g)
{
    using vec_t = typename SK<T>::vec_t;
    using B_ = Blk<T, lines_of<T>(L_KC, MT)>;
    constexpr int V = B_::V;
    // ---- L ----
    if (L_KC) {
        // column tile t: row n0 + 16t + c, vector q: k_of(k0, g, qV) .. + V-1
#pragma unroll
        for (int t = 0; t < 4; ++t) {
#pragma unroll
            for (int q = 0; q < B_::RV; ++q) {
                const int64_t row = n0 + 16 * t + c;
                const int64_t kk = k_of<T, true>(k0, g, q * V);
                vec_t v;
                if (a.l_vec && row < a.Nl && kk + V <= kend) {
                    v = *reinterpret_cast<const vec_t *>(a.L + row * a.ldl + kk);
                } else {
#pragma unroll
                    for (int e = 0; e < V; ++e)
                        v[e] = (row < a.Nl && kk + e < kend) ? a.L[row * a.ldl + kk + e] : T(0);
                }
                b.l[t * B_::RV + q] = v;
            }
        }
    } else {
        // step j: row k_of(k0, g, j), columns n0 + h*16V + V*c .. + V-1
#pragma unroll
        for (int j = 0; j < B_::NSTEP; ++j) {
#pragma unroll
            for (int h = 0; h < 4 / V; ++h) {
                const int64_t kk = k_of<T, false>(k0, g, j);
                const int64_t col = n0 + h * 16 * V + V * c;
                vec_t v;
                if (a.l_vec && kk < kend && col + V <= a.Nl) {
                    v = stream_load(reinterpret_cast<const vec_t *>(a.L + kk * a.ldl + col));
                } else {
#pragma unroll
                    for (int e = 0; e < V; ++e)
                        v[e] = (kk < kend && col + e < a.Nl) ? a.L[kk * a.ldl + col + e] : T(0);
                }
                b.l[j * (4 / V) + h] = v;
            }
        }
    }
    // ---- S ----
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        const int64_t ms = mt * 16 + c;
        if (SM != 0) {
#pragma unroll
            for (int q = 0; q < B_::RV; ++q) {
                const int64_t kk = k_of<T, L_KC>(k0, g, q * V);
                vec_t v;
                if (SM == 2 && ms < a.Ms && kk + V <= kend) {
                    v = *reinterpret_cast<const vec_t *>(a.S + ms * a.lds + kk);
                } else {
#pragma unroll
                    for (int e = 0; e < V; ++e)
                        v[e] = (ms < a.Ms && kk + e < kend) ? a.S[ms * a.lds + kk + e] : T(0);
                }
                b.s_run[mt][q] = v;
            }
        } else {
#pragma unroll
            for (int j = 0; j < B_::NSTEP; ++j) {
                const int64_t kk = k_of<T, L_KC>(k0, g, j);
                b.s_scl[mt][j] = (ms < a.Ms && kk < kend) ? a.S[kk * a.lds + ms] : T(0);
            }
        }
    }
}

// The staging buffer: one K-contiguous block row per O column, 16-byte chunk
// ch of row r in slot ch ^ stage_swz(r), so that 32 consecutive lanes of the
// writes (4 or 2 whole rows) and of the operand reads (rows 16 t + c, c < 16,
// chunks 4 q + g for two lane groups g) fall on 32 distinct 16-byte bank
// groups of 512 bytes.  Measured on MI355X: rows of one 128-byte line (FP32,
// and FP64 at 32 rows) went from 2 SQ_LDS_BANK_CONFLICT cycles per staging
// access (rows padded by 16 bytes) to none, the 32-row K-contiguous cases
// 0.8 % faster on average; FP64's two-line rows (8 and 16 rows, already at
// 6.5-6.7 TB/s) still count 2.
template <typename T, int MT>
struct Stage {
    static constexpr int ROW = Blk<T, lines_of<T>(true, MT)>::KB;
    static constexpr int ELEMS = kPanel * ROW;   // per wave
};

template <int CPR>
__device__ __forceinline__ int stage_swz(int r)
{
    static_assert(CPR == 8 || CPR == 16, "one or two 128-byte lines per row");
    return CPR == 8 ? 2 * (((r >> 1) ^ (r >> 3)) & 3) : 2 * ((r >> 1) & 7);
}

// Staged K-contiguous L: write the lane's raw line pieces to the wave's own
// LDS rows, then read back the MFMA operands (row 16 t + c, k_of(k0, g, qV)
// .. + V-1).  Only this wave touches its rows and LDS operations of one wave
// execute in order, so no barrier is needed; the wave barrier keeps the
// compiler from moving the reads above the writes.
template <typename T, int MT, bool S_KC>
__device__ __forceinline__ void stage_block(Block<T, MT, S_KC, true> &b, T *w, int c, int g)
{
    using vec_t = typename SK<T>::vec_t;
    using B_ = Blk<T, lines_of<T>(true, MT)>;
    constexpr int V = B_::V, ROW = Stage<T, MT>::ROW, CPR = B_::CPR;
    const int lane = g * 16 + c;
#pragma unroll
    for (int i = 0; i < B_::NLV; ++i) {
        const int r = B_::RPI * i + lane / CPR;
        *reinterpret_cast<vec_t *>(w + r * ROW + (((lane % CPR) ^ stage_swz<CPR>(r)) * V)) = b.l[i];
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int q = 0; q < B_::RV; ++q) {
            const int r = 16 * t + c;
            b.l[t * B_::RV + q] = *reinterpret_cast<const vec_t *>(w + r * ROW + (((4 * q + g) ^ stage_swz<CPR>(r)) * V));
        }
    __builtin_amdgcn_wave_barrier();
}

template <typename T, int MT, bool S_KC, bool L_KC>
__device__ __forceinline__ void mma_block(typename SK<T>::acc_t (&acc)[MT][4],
                                          const Block<T, MT, S_KC, L_KC> &b)
{
#pragma unroll
    for (int j = 0; j < Blk<T, lines_of<T>(L_KC, MT)>::NSTEP; ++j)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[mt][t] = SK<T>::mma(b.sop(mt, j), b.lop(t, j), acc[mt][t]);
}

// One step of the ring: multiply block b.  The scheduling barrier keeps the
// loads of later steps below this step's MFMAs: a ring slot refilled by a
// later step is a new value to the compiler, and hoisting its loads would
// keep one more block live than the ring holds.
template <typename T, int MT, bool S_KC, bool L_KC>
__device__ __forceinline__ void consume_block(typename SK<T>::acc_t (&acc)[MT][4], Block<T, MT, S_KC, L_KC> &b,
                                              T *w, int c, int g)
{
    if constexpr (L_KC) stage_block<T, MT, S_KC>(b, w, c, g);
    mma_block<T, MT, S_KC, L_KC>(acc, b);
    __builtin_amdgcn_sched_barrier(0);
}

// Waves per SIMD the register allocation must allow, and DEPTH: K blocks
// whose loads are in flight while a wave multiplies the current one.  Resident
// waves times DEPTH sets the bytes in flight per CU.  Measured on MI355X at
// K = 8192, long side 65536 (all modes): one wave per SIMD with a ring of 4
// (16 rows) or 3 (32 rows) blocks streams as fast as or up to 4.5 % faster
// than 2-4 waves with rings of 1-3, and ring depths 3-8 are alike.  One wave
// per SIMD also needs no K split at that size.  64-row FP32 keeps 2 waves, 1
// block.
// With K-split waves (KS = 2) each wave has half the registers, so the rings
// are shorter: 8 waves x 3 (16 rows) or x 2 (32 rows) blocks per CU.
template <typename T, int MT, bool L_KC>
struct SkinnyOcc {
    static constexpr int KS = skinny_ks<T>(MT, L_KC);
    static constexpr int WPE = KS * (MT <= 2 ? 1 : (sizeof(T) == 8 ? 1 : 2));
    // FP64 blocks of K-contiguous L are kLinesKC times larger
    static constexpr int D = KS == 2 ? (MT == 1 ? 3 : 2)
                                     : (lines_of<T>(L_KC, MT) == 2 ? 2 : (MT == 1 ? 4 : (MT == 2 ? 3 : 1)));
};

// LDS of one workgroup: the staging rows of K-contiguous L (one set per wave)
// and, with K-split waves, the partial sums of the second half (MT x 4 tiles
// x 4 registers x 64 lanes per panel), which reuse the staging rows.
template <typename T, int MT, bool L_KC>
constexpr int skinny_lds_elems()
{
    constexpr int KS = skinny_ks<T>(MT, L_KC);
    constexpr int st = L_KC ? KS * kWaves * Stage<T, MT>::ELEMS : 0;
    constexpr int red = KS > 1 ? kWaves * MT * 4 * 4 * 64 : 0;
    return st > red ? st : (red > 0 ? red : 1);
}

// f(std::integral_constant<int, 0>{}) .. f(std::integral_constant<int, N - 1>{})
template <int N, int I = 0, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<N, I + 1>(f);
    }
}

// Shared S in chunks (16 or 32 rows: MT = 1 with Ms > 8, or MT = 2; k-major
// L; S 16-byte aligned with 16-byte aligned rows and, stored m-contiguous, Ms
// a multiple of the vector).  Every wave of a workgroup (of one K part)
// multiplies the same S block, which at 32 rows is half the size of the
// wave's L block (a quarter at 16); read by each wave, it puts that many
// bytes again through the memory pipeline.  Measured on MI355X at K = 8192,
// long side 65536: pointing every lane's S load at one address (a
// timing-only copy) ran the k-major cases 6-14 % faster at 32 rows, 2-11 % at
// 16 and not measurably at 8.  Shared through LDS block by block, behind a
// barrier per block, the 32-row cases ran 1-9 % slower: with one workgroup
// per CU every wave then waits for the slowest one's loads at every block.
// So here the workgroup shares S in chunks of SHB blocks (two ring rounds)
// with one barrier per chunk:
//   * the chunk (ROWS = 16 MT rows x CHK k) sits in LDS as [row][k] (rows
//     padded by 16 bytes, so the 16-byte operand reads of 8 consecutive rows
//     hit distinct banks), two slots per K part; S stored m-contiguous is
//     read as V rows at one k per 16 bytes and written to LDS element by
//     element;
//   * every lane stages STG 16-byte vectors of each chunk, one at every
//     EVERY-th block (every block at 32 rows, every other at 16): at block b
//     of chunk cc (position p = b % SHB, p a multiple of EVERY) a wave writes
//     its staged vector p / EVERY of chunk cc + 1 to slot (cc + 1) & 1 and
//     loads the same vector of chunk cc + 2 into that register; at every
//     block it reads its S operands from slot cc & 1;
//   * at p = 0 a barrier: every wave has finished chunk cc - 1, so slot
//     (cc + 1) & 1 = (cc - 1) & 1 is free and slot cc & 1 is complete.
// The loads at each chunk position are fixed, so the wait before each
// block's MFMAs still counts a known number of loads.  Measured (interleaved
// A/B, bit-identical), 32 rows: S K-contiguous (shortA NN, thinB TT) +10-11 %
// FP32 and +8.5-9 % FP64; S m-contiguous (shortA TN, thinB TN) +4.7 % FP32 and
// +1-6 % FP64; 16 rows, the same eight cases: +2-8 %.
template <typename T, int MT, int SM, bool L_KC>
__host__ __device__ constexpr bool skinny_share()
{
    return MT <= 2 && !L_KC && SM != 1;
}

template <typename T, int MT>
struct ShChunk {
    using B_ = Blk<T, 1>;
    static constexpr int V = B_::V, KB = B_::KB, RUN = B_::RUN, RV = B_::RV;
    static constexpr int ROWS = 16 * MT;
    static constexpr int D = SkinnyOcc<T, MT, false>::D;
    static constexpr int SHB = 2 * (D + 1);              // blocks per chunk: whole ring rounds
    static constexpr int CHK = SHB * KB;                 // k per chunk
    static constexpr int ROW = CHK + 16 / (int)sizeof(T);
    static constexpr int SLOT = ROWS * ROW;              // elements per slot
    static constexpr int VPR = CHK / V;                  // 16-byte vectors per chunk row
    static constexpr int STG = ROWS * VPR / (kWaves * 64);  // vectors per lane per chunk
    static constexpr int EVERY = SHB / STG;              // blocks per staged vector
    static_assert(STG * kWaves * 64 == ROWS * VPR && STG * EVERY == SHB, "whole vectors per lane");
};

// Vector v of the chunk at kc.  S K-contiguous (SM 2): row v / VPR (rows
// past Ms read row Ms - 1), k = kc + (v % VPR) V.  S m-contiguous (SM 0):
// k = kc + v / (ROWS / V), rows (v % (ROWS / V)) V .. + V - 1 (a vector
// wholly past Ms reads rows from 0; Ms is a multiple of V).  Past kend (the
// end of the shared blocks, whose data is never used) it reads k = kc
// instead, which is inside the range.
template <typename T, int MT, int SM>
__device__ __forceinline__ typename SK<T>::vec_t sh_load(const SkinnyArgs<T> &a, int64_t kc, int64_t kend, int v)
{
    using C_ = ShChunk<T, MT>;
    constexpr int64_t s = sizeof(T);
    if constexpr (SM == 2) {
        const int row = v / C_::VPR, col = v % C_::VPR;
        const int64_t r = row < a.Ms ? row : a.Ms - 1;
        const int64_t kk = kc + (int64_t)col * C_::V + C_::V <= kend ? (int64_t)col * C_::V : 0;
        return bload_vec<T, 0>(a.S + kc, (uint32_t)((r * a.lds + kk) * s));
    } else {
        const int kk = v / (C_::ROWS / C_::V), r0 = (v % (C_::ROWS / C_::V)) * C_::V;
        const int64_t r = r0 < a.Ms ? r0 : 0;
        const int64_t kq = kc + kk < kend ? kk : 0;
        return bload_vec<T, 0>(a.S + kc * a.lds, (uint32_t)((kq * a.lds + r) * s));
    }
}
template <typename T, int MT, int SM>
__device__ __forceinline__ void sh_store(T *slot, int v, typename SK<T>::vec_t x)
{
    using C_ = ShChunk<T, MT>;
    if constexpr (SM == 2) {
        *reinterpret_cast<typename SK<T>::vec_t *>(slot + (v / C_::VPR) * C_::ROW + (v % C_::VPR) * C_::V) = x;
    } else {
        const int kk = v / (C_::ROWS / C_::V), r0 = (v % (C_::ROWS / C_::V)) * C_::V;
#pragma unroll
        for (int e = 0; e < C_::V; ++e) slot[(r0 + e) * C_::ROW + kk] = x[e];
    }
}

template <typename T, int MT, int SM, bool L_KC, bool TRANS_OUT>
__global__ void __launch_bounds__(256 * skinny_ks<T>(MT, L_KC))
__attribute__((amdgpu_waves_per_eu(SkinnyOcc<T, MT, L_KC>::WPE, SkinnyOcc<T, MT, L_KC>::WPE)))
skinny_kernel(SkinnyArgs<T> a)
{
    constexpr bool S_KC = SM != 0;
    constexpr int KS = skinny_ks<T>(MT, L_KC);
    using acc_t = typename SK<T>::acc_t;
    // the wave index is made wave-uniform, so n0 and the load bases below
    // live in scalar registers
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c = lane & 15, g = lane >> 4;
    const int pw = wave % kWaves, half = wave / kWaves;  // panel; part of K (KS = 2)
    const int64_t n0 = ((int64_t)blockIdx.x * kWaves + pw) * kPanel;
    const bool active = n0 < a.Nl;
    // with KS = 1 there are no barriers in this kernel: idle waves may leave
    if (KS == 1 && !active) return;
    const int split = blockIdx.y;
    int64_t kb = (int64_t)split * a.kslice;
    int64_t ke = min(a.K, kb + a.kslice);

    constexpr int KB = Blk<T, lines_of<T>(L_KC, MT)>::KB;
    constexpr bool SHARE = skinny_share<T, MT, SM, L_KC>();
    // full K blocks that every wave of the workgroup walks in step (shared S)
    int64_t nbs = 0;
    if (KS == 2) {
        // the first wave of a panel takes the first floor(blocks / 2) K
        // blocks, the second the rest (with the ragged block, if any)
        const int64_t kmid = kb + ((ke - kb + KB - 1) / KB / 2) * KB;
        if (SHARE) nbs = min((kmid - kb) / KB, (ke - kmid) / KB);
        if (half == 0) ke = kmid;
        else kb = kmid;
    } else if (SHARE) {
        nbs = (ke - kb) / KB;
    }
    // Shared S needs every panel of the workgroup full, so that every wave
    // walks nbs blocks and meets the same barriers; workgroup-uniform.
    const bool shared = SHARE && nbs > 0 && a.s_vec && a.l_vec && (SM != 0 || a.Ms % SK<T>::V == 0) &&
                        (MT == 2 || a.Ms > 8) && ((int64_t)blockIdx.x + 1) * kWaves * kPanel <= a.Nl;

    acc_t acc[MT][4];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[i][t] = acc_t{0, 0, 0, 0};

    const bool panel_full = a.l_vec && (n0 + kPanel <= a.Nl);
    const int64_t kfull = panel_full ? kb + ((ke - kb) / KB) * KB : kb;

    // K-contiguous L is staged through the wave's own LDS rows.
    __shared__ __attribute__((aligned(16))) T stage[skinny_lds_elems<T, MT, L_KC>()];
    T *const w = stage + (L_KC ? wave * Stage<T, MT>::ELEMS : 0);
    constexpr int SH_MT = SHARE ? MT : 1;  // chunk geometry exists for MT 1 and 2 only
    __shared__ __attribute__((aligned(16))) T sshare[SHARE ? KS * 2 * ShChunk<T, SH_MT>::SLOT : 1];
    int64_t k0 = kfull;
    if (active) {
        // Full blocks: unmasked loads into a ring of D + 1 register blocks.
        // Block j lives in b[j % (D + 1)]; while block j is multiplied, the loads
        // of blocks j + 1 .. j + D are in flight.  The steady-state loop runs
        // whole rounds of D + 1 blocks, so every ring index is a compile-time
        // constant and no block is copied between registers.
        constexpr int D = SkinnyOcc<T, MT, L_KC>::D;
        if constexpr (SHARE) {
            if (shared) {
                using C_ = ShChunk<T, SH_MT>;
                using vec_t = typename SK<T>::vec_t;
                constexpr int SHB = C_::SHB, STG = C_::STG, EVERY = C_::EVERY;
                static_assert(C_::D == D, "a chunk is two ring rounds");
                const LaneOff<MT> o = lane_offsets<T, MT, SM, L_KC>(a, c, g);
                T *const sh = sshare + half * (2 * C_::SLOT);
                const int64_t kend = kb + nbs * KB, nch = (nbs + SHB - 1) / SHB;
                // chunk start; chunks past the last reload chunk 0 (never used)
                auto chunk_k = [&](int64_t cc) { return cc < nch ? kb + cc * C_::CHK : kb; };
                const int vb = pw * 64 + lane;  // this lane's vector q is q * 256 + vb
                vec_t stg[STG];
                // chunk 0 to slot 0, chunk 1 to the staging registers; block
                // 0's barrier orders the slot writes before any read
#pragma unroll
                for (int q = 0; q < STG; ++q) stg[q] = sh_load<T, MT, SM>(a, chunk_k(0), kend, q * 256 + vb);
#pragma unroll
                for (int q = 0; q < STG; ++q) sh_store<T, MT, SM>(sh, q * 256 + vb, stg[q]);
#pragma unroll
                for (int q = 0; q < STG; ++q) stg[q] = sh_load<T, MT, SM>(a, chunk_k(1), kend, q * 256 + vb);
                // block j at chunk position p = j % SHB (a compile-time value)
                auto use = [&](Block<T, MT, S_KC, L_KC> &x, int64_t j, auto pc) {
                    constexpr int p = decltype(pc)::value;
                    const int64_t cc = j / SHB;
                    if constexpr (p == 0) __syncthreads();
                    if constexpr (p % EVERY == 0) {
                        constexpr int q = p / EVERY;
                        sh_store<T, MT, SM>(sh + ((cc + 1) & 1) * C_::SLOT, q * 256 + vb, stg[q]);
                        stg[q] = sh_load<T, MT, SM>(a, chunk_k(cc + 2), kend, q * 256 + vb);
                    }
                    const T *src = sh + (cc & 1) * C_::SLOT + c * C_::ROW + p * KB + g * C_::RUN;
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                        for (int q = 0; q < C_::RV; ++q) {
                            const vec_t sv = *reinterpret_cast<const vec_t *>(src + mt * 16 * C_::ROW + q * C_::V);
                            if constexpr (S_KC) {
                                x.s_run[mt][q] = sv;
                            } else {
#pragma unroll
                                for (int e = 0; e < C_::V; ++e) x.s_scl[mt][q * C_::V + e] = sv[e];
                            }
                        }
                    mma_block<T, MT, S_KC, L_KC>(acc, x);
                    __builtin_amdgcn_sched_barrier(0);
                };
                Block<T, MT, S_KC, L_KC> b[D + 1];
                int64_t i = 0;
                if (nbs >= SHB + D) {
                    // unconditional first loads: see the ring below
#pragma unroll
                    for (int u = 0; u < D; ++u) {
                        load_full<T, MT, SM, L_KC, false>(b[u], a, o, n0, kb + u * KB);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                    for (; i + SHB + D <= nbs; i += SHB) {
                        static_for<SHB>([&](auto uc) {
                            constexpr int u = decltype(uc)::value;
                            load_full<T, MT, SM, L_KC, false>(b[(u + D) % (D + 1)], a, o, n0,
                                                              kb + (i + u + D) * KB);
                            use(b[u % (D + 1)], i + u, uc);
                        });
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < D; ++u)
                        if (u < nbs) load_full<T, MT, SM, L_KC, false>(b[u], a, o, n0, kb + u * KB);
                }
                // the last nbs - i < SHB + D blocks; blocks i .. i + D - 1 are loaded
                static_for<SHB + D>([&](auto uc) {
                    constexpr int u = decltype(uc)::value;
                    if (i + u < nbs) {
                        if (i + u + D < nbs)
                            load_full<T, MT, SM, L_KC, false>(b[(u + D) % (D + 1)], a, o, n0, kb + (i + u + D) * KB);
                        use(b[u % (D + 1)], i + u, std::integral_constant<int, u % SHB>{});
                    }
                });
                k0 = kend;
            }
        }
        const int64_t nb = shared ? 0 : (kfull - kb) / KB;
        if (nb > 0) {
            const LaneOff<MT> o = lane_offsets<T, MT, SM, L_KC>(a, c, g);
            Block<T, MT, S_KC, L_KC> b[D + 1];
            int64_t i = 0;
            if (nb > 2 * D) {
                // The first D blocks are loaded unconditionally and in ring
                // order (a scheduling barrier after each) on the way into the
                // steady loop.  The wait before each block's MFMAs is counted
                // at compile time, and at the loop head the compiler takes the
                // fewest loads issued after a block on any incoming path;
                // with the first blocks' loads reordered or conditional, the
                // first block of every round waited for all but the newest
                // block in flight (the other resident waves covered it: the
                // fix measured within noise at K = 8192).
#pragma unroll
                for (int u = 0; u < D; ++u) {
                    load_full<T, MT, SM, L_KC>(b[u], a, o, n0, kb + u * KB);
                    __builtin_amdgcn_sched_barrier(0);
                }
                for (; i + 2 * D < nb; i += D + 1) {
#pragma unroll
                    for (int u = 0; u <= D; ++u) {
                        load_full<T, MT, SM, L_KC>(b[(u + D) % (D + 1)], a, o, n0, kb + (i + u + D) * KB);
                        consume_block<T, MT, S_KC, L_KC>(acc, b[u], w, c, g);
                    }
                }
            } else {
#pragma unroll
                for (int u = 0; u < D; ++u)
                    if (u < nb) load_full<T, MT, SM, L_KC>(b[u], a, o, n0, kb + u * KB);
            }
            // the last nb - i <= 2D blocks; blocks i .. i + D - 1 are loaded
#pragma unroll
            for (int u = 0; u < 2 * D; ++u) {
                if (i + u < nb) {
                    if (i + u + D < nb)
                        load_full<T, MT, SM, L_KC>(b[(u + D) % (D + 1)], a, o, n0, kb + (i + u + D) * KB);
                    consume_block<T, MT, S_KC, L_KC>(acc, b[u % (D + 1)], w, c, g);
                }
            }
        }
        // Ragged tail (partial panel or partial K block, and in shared mode
        // the blocks past the shared ones): masked loads.
        for (; k0 < ke; k0 += KB) {
            Block<T, MT, S_KC, L_KC> b;
            load_tail<T, MT, SM, L_KC>(b, a, n0, k0, ke, c, g);
            mma_block<T, MT, S_KC, L_KC>(acc, b);
        }
    }  // active

    if constexpr (KS == 2) {
        // The second wave of each panel hands its sums to the first, which
        // adds them (first part + second part, a fixed order) and stores.
        // Every wave of the workgroup reaches both barriers.
        __syncthreads();  // all staging reads are done: the rows are reused
        T *red = stage + pw * (MT * 4 * 4 * 64);
        if (half == 1 && active) {
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int r = 0; r < 4; ++r) red[((mt * 4 + t) * 4 + r) * 64 + lane] = acc[mt][t][r];
        }
        __syncthreads();
        if (half == 1 || !active) return;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[mt][t][r] += red[((mt * 4 + t) * 4 + r) * 64 + lane];
    }

    // Store: lane (c, g), register r of tile (mt, t) is
    // O[mt*16 + out_row(g, r)][n0 + col_of(t, c)].
    auto put = [&](int mt, int t, int r) {
        const int64_t nl = n0 + col_of<T, L_KC>(t, c);
        const int64_t ms = mt * 16 + SK<T>::out_row(g, r);
        if (nl >= a.Nl || ms >= a.Ms) return;
        const T v = acc[mt][t][r];
        if (a.splitk > 1) {
            a.W[((int64_t)split * a.Ms + ms) * a.Nl + nl] = v;
        } else {
            T *p = TRANS_OUT ? a.C + nl * a.ldc + ms : a.C + ms * a.ldc + nl;
            T o = a.alpha * v;
            if (a.beta != T(0)) o += a.beta * *p;
            *p = o;
        }
    };
    // V consecutive elements of one row of C held by one lane go out as one
    // 16-byte access (the same arithmetic per element): element-wise, each
    // store instruction would write 64 scattered words, and every wave reaches
    // its stores at about the same time at the end of the K walk.  Measured on
    // MI355X at K = 8192, long side 65536 (interleaved A/B, bit-identical): the
    // FP32 thinB cases +0.1-2.2 %, FP32 shortA with k-major L +0.2-3.5 %, FP64
    // unchanged within noise.
    using vec_t = typename SK<T>::vec_t;
    constexpr int V = SK<T>::V;
    const bool vec_c = a.c_vec && a.splitk == 1;
    auto put_vec = [&](T *p, vec_t v) {
        vec_t o = a.alpha * v;
        if (a.beta != T(0)) {
            const vec_t old = *reinterpret_cast<const vec_t *>(p);
#pragma unroll
            for (int e = 0; e < V; ++e) o[e] += a.beta * old[e];
        }
        *reinterpret_cast<vec_t *>(p) = o;
    };
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        if constexpr (TRANS_OUT && sizeof(T) == 4) {
            // FP32 output rows 4 g + r: registers r = 0..3 of tile (mt, t) are
            // C[nl][mt*16 + 4g .. + 3]
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int64_t nl = n0 + col_of<T, L_KC>(t, c), ms0 = mt * 16 + 4 * g;
                if (vec_c && nl < a.Nl && ms0 + 4 <= a.Ms) {
                    put_vec(a.C + nl * a.ldc + ms0, acc[mt][t]);
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r) put(mt, t, r);
                }
            }
        } else if constexpr (!TRANS_OUT && !L_KC) {
            // k-major L: tiles h V + e hold columns n0 + h 16 V + V c + e, so
            // register r of those V tiles is C[ms][n0 + h 16 V + V c .. + V - 1]
#pragma unroll
            for (int r = 0; r < 4; ++r) {
#pragma unroll
                for (int h = 0; h < 4 / V; ++h) {
                    const int64_t ms = mt * 16 + SK<T>::out_row(g, r);
                    const int64_t nl0 = n0 + h * 16 * V + V * c;
                    if (vec_c && ms < a.Ms && nl0 + V <= a.Nl) {
                        vec_t v;
#pragma unroll
                        for (int e = 0; e < V; ++e) v[e] = acc[mt][h * V + e][r];
                        put_vec(a.C + ms * a.ldc + nl0, v);
                    } else {
#pragma unroll
                        for (int e = 0; e < V; ++e) put(mt, h * V + e, r);
                    }
                }
            }
        } else {
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) put(mt, t, r);
        }
    }
}

// Sum the K slices in order s = 0, 1, ... and apply the epilogue.
template <typename T, bool TRANS_OUT>
__global__ void skinny_reduce_kernel(SkinnyArgs<T> a)
{
    const int64_t total = a.Ms * a.Nl;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        // TRANS_OUT walks C row-major (nl outer) so the stores coalesce.
        const int64_t ms = TRANS_OUT ? e % a.Ms : e / a.Nl;
        const int64_t nl = TRANS_OUT ? e / a.Ms : e % a.Nl;
        T sum = T(0);
        for (int s = 0; s < a.splitk; ++s) sum += a.W[((int64_t)s * a.Ms + ms) * a.Nl + nl];
        T *p = TRANS_OUT ? a.C + nl * a.ldc + ms : a.C + ms * a.ldc + nl;
        T o = a.alpha * sum;
        if (a.beta != T(0)) o += a.beta * *p;
        *p = o;
    }
}

template <typename T, int MT, int SM, bool L_KC, bool TRANS_OUT>
hipError_t launch4(const SkinnyArgs<T> &a, hipStream_t s)
{
    const int64_t wgs = (a.Nl + kWaves * kPanel - 1) / (kWaves * kPanel);
    hipLaunchKernelGGL((skinny_kernel<T, MT, SM, L_KC, TRANS_OUT>), dim3((unsigned)wgs, a.splitk),
                       dim3(kWaves * 64 * skinny_ks<T>(MT, L_KC)), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || a.splitk == 1) return e;
    const int64_t total = a.Ms * a.Nl;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL((skinny_reduce_kernel<T, TRANS_OUT>), dim3((unsigned)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

template <typename T, int MT, int SM>
hipError_t launch_sm(const SkinnyArgs<T> &a, bool l_kc, bool trans_out, hipStream_t s)
{
    if (l_kc) return trans_out ? launch4<T, MT, SM, true, true>(a, s) : launch4<T, MT, SM, true, false>(a, s);
    return trans_out ? launch4<T, MT, SM, false, true>(a, s) : launch4<T, MT, SM, false, false>(a, s);
}

template <typename T, int MT>
hipError_t launch_mt(const SkinnyArgs<T> &a, bool s_kc, bool l_kc, bool trans_out, hipStream_t s)
{
    if (!s_kc) return launch_sm<T, MT, 0>(a, l_kc, trans_out, s);
    if (!a.s_vec) return launch_sm<T, MT, 1>(a, l_kc, trans_out, s);
    return launch_sm<T, MT, 2>(a, l_kc, trans_out, s);
}

// Operand roles for a validated GEMM.  shortA when M <= N, thinB otherwise.
template <typename T>
void roles(const Gemm<T> &g, SkinnyArgs<T> &a, bool &s_kc, bool &l_kc, bool &trans_out)
{
    constexpr int V = SK<T>::V;
    if (g.M <= g.N) {  // shortA: S = op(A), L = op(B)
        a.S = g.A; a.lds = g.lda; a.L = g.B; a.ldl = g.ldb;
        a.Ms = g.M; a.Nl = g.N;
        s_kc = !g.ta;    // A stored M x K (k contiguous) unless transposed
        l_kc = g.tb;     // B stored N x K (k contiguous) when transposed
        trans_out = false;
    } else {           // thinB: S = op(B)^T, L = op(A)^T, O = C^T
        a.S = g.B; a.lds = g.ldb; a.L = g.A; a.ldl = g.lda;
        a.Ms = g.N; a.Nl = g.M;
        s_kc = g.tb;     // op(B)^T[n][k] = B[n][k] when B is stored N x K
        l_kc = !g.ta;    // op(A)^T[k][m] = A[m][k]: A stored M x K is m-major rows of k
        trans_out = true;
    }
    a.K = g.K;
    a.C = g.C; a.ldc = g.ldc;
    a.alpha = g.alpha; a.beta = g.beta;
    a.l_vec = ((uintptr_t)a.L % 16 == 0) && (a.ldl % V == 0);
    a.s_vec = ((uintptr_t)a.S % 16 == 0) && (a.lds % V == 0);
    a.c_vec = ((uintptr_t)a.C % 16 == 0) && (a.ldc % V == 0);
}

}  // namespace

template <typename T>
bool skinny_fits(const Gemm<T> &g)
{
    SkinnyArgs<T> a{};
    bool s_kc, l_kc, trans_out;
    roles(g, a, s_kc, l_kc, trans_out);
    if (a.Ms > 64 || a.Ms <= 0 || a.Nl <= 0) return false;
    // largest lane offsets of lane_offsets(), in bytes (rows of S up to 63)
    constexpr double V = SK<T>::V, RUN = Blk<T, 1>::RUN, s = sizeof(T);
    const int lines = lines_of<T>(l_kc, a.Ms <= 16 ? 1 : (a.Ms <= 32 ? 2 : (a.Ms <= 48 ? 3 : 4)));
    const double CPR = lines * 128 / s / V, RPI = 64 / CPR;
    const double kg = 3 * (l_kc ? V : RUN);
    const double lmax = l_kc ? ((RPI - 1) * (double)a.ldl + (CPR - 1) * V) * s : (3 * RUN * (double)a.ldl + 15 * V) * s;
    // (the shared chunk loads reach up to 512 elements along k)
    const double smax = s_kc ? (63.0 * (double)a.lds + (kg > 512 ? kg : 512)) * s
                             : ((kg > 512 ? kg : 512) * (double)a.lds + 63) * s;
    // a 16-byte access at the offset stays inside the buffer resource's 2^31 - 1 bytes
    return lmax + 16 < 2147483647.0 && smax + 16 < 2147483647.0;
}

// Whether the skinny kernel can occupy every resident workgroup slot: its
// panels times the most K slices it may take (skinny_pick_splitk) reach one
// full round.  Below that the tile kernel with 64 tiles and split-K runs the
// shape faster: measured on MI355X over 32 shapes of a mixed GEMM workload,
// long sides of a few hundred to a few thousand ran 1.1-4.5x slower on the
// skinny kernel, and the autotuner never picked it there.
template <typename T>
bool skinny_fills(const Gemm<T> &g, int num_cus)
{
    const int64_t ms = g.M <= g.N ? g.M : g.N;
    const int64_t nl = g.M <= g.N ? g.N : g.M;
    const int wg_cu = ms <= 16 ? SkinnyOcc<T, 1, false>::WPE / SkinnyOcc<T, 1, false>::KS
                               : (ms <= 32 ? SkinnyOcc<T, 2, false>::WPE / SkinnyOcc<T, 2, false>::KS
                                           : SkinnyOcc<T, 4, false>::WPE / SkinnyOcc<T, 4, false>::KS);
    const int64_t wgs = (nl + kWaves * kPanel - 1) / (kWaves * kPanel);
    const int64_t slots = (int64_t)wg_cu * (num_cus > 0 ? num_cus : 256);
    int64_t max_s = g.K / 512;
    if (max_s > 64) max_s = 64;
    if (max_s < 1) max_s = 1;
    return wgs * max_s >= slots;
}

template <typename T>
int skinny_pick_splitk(const Gemm<T> &g, int num_cus)
{
    // Cost of a launch in units of one full round of unsplit panel walks over
    // K: s slices take ceil(wgs * s / slots) rounds of 1/s each (a round that
    // is not full leaves CUs idle), plus the partial sums, written once per
    // slice and read once by the reduction, against what a full round
    // streams (slots panels' K x 256 block of L).  Fewer slices win
    // near-ties.  Pricing the partial sums against the streamed K x Nl
    // instead overcharged them by slots / wgs: on FP64 3024 x 20480 x 48 it
    // took 12 slices at 17.9 TFLOP/s where 21 (one full round) run 29.4, and
    // FP32 3024 x 20480 x 16 took 18 slices at 29.2 where 20 run 31.3.
    const int64_t ms = g.M <= g.N ? g.M : g.N;
    const int64_t nl = g.M <= g.N ? g.N : g.M;
    // resident workgroups per CU: waves per SIMD over the K-split waves per
    // panel (the same for either layout of L)
    const int wg_cu = ms <= 16 ? SkinnyOcc<T, 1, false>::WPE / SkinnyOcc<T, 1, false>::KS
                               : (ms <= 32 ? SkinnyOcc<T, 2, false>::WPE / SkinnyOcc<T, 2, false>::KS
                                           : SkinnyOcc<T, 4, false>::WPE / SkinnyOcc<T, 4, false>::KS);
    const int64_t wgs = (nl + kWaves * kPanel - 1) / (kWaves * kPanel);
    const int64_t slots = (int64_t)wg_cu * (num_cus > 0 ? num_cus : 256);
    int64_t max_s = g.K / 512;  // slices of >= 512
    if (max_s > 64) max_s = 64;
    int best = 1;
    const double fill = 256.0 / (double)g.K;  // a walk's fixed cost: ring fill and drain, epilogue
    double best_cost = (double)((wgs + slots - 1) / slots) * (1.0 + fill);
    const double per_slice = (double)ms * (double)nl / ((double)g.K * (double)slots * (kWaves * kPanel));
    for (int64_t sk = 2; sk <= max_s; ++sk) {
        const double rounds = (double)((wgs * sk + slots - 1) / slots);
        const double cost = rounds * (1.0 / (double)sk + fill) + (double)(2 * sk + 1) * per_slice;
        if (cost < 0.97 * best_cost) {
            best = (int)sk;
            best_cost = cost;
        }
    }
    return best;
}

template <typename T>
hipError_t launch_skinny(const Gemm<T> &g, int splitk, T *work, hipStream_t s)
{
    if (!skinny_fits(g)) return hipErrorInvalidValue;
    SkinnyArgs<T> a{};
    bool s_kc, l_kc, trans_out;
    roles(g, a, s_kc, l_kc, trans_out);
    if (splitk < 1) splitk = 1;
    const int KB = lines_of<T>(l_kc, a.Ms <= 16 ? 1 : (a.Ms <= 32 ? 2 : (a.Ms <= 48 ? 3 : 4))) * Blk<T, 1>::KB;
    int64_t kslice = (g.K + splitk - 1) / splitk;
    kslice = (kslice + KB - 1) / KB * KB;
    const int64_t nsplit = (g.K + kslice - 1) / kslice;
    a.kslice = kslice;
    a.splitk = (int)(nsplit < 1 ? 1 : nsplit);
    if (a.splitk > 1 && work == nullptr) return hipErrorInvalidValue;
    a.W = work;
    if (a.Ms <= 16) return launch_mt<T, 1>(a, s_kc, l_kc, trans_out, s);
    if (a.Ms <= 32) return launch_mt<T, 2>(a, s_kc, l_kc, trans_out, s);
    // 48 rows as three row tiles rather than 64 padded: a quarter fewer
    // MFMAs and accumulators; at 40 and 48 rows (both precisions, all modes,
    // long sides 3024-65536, K 2048 and 8192) 1.10-1.30x faster, median 1.2x,
    // with identical results.
    if (a.Ms <= 48) return launch_mt<T, 3>(a, s_kc, l_kc, trans_out, s);
    return launch_mt<T, 4>(a, s_kc, l_kc, trans_out, s);
}

template bool skinny_fits<double>(const Gemm<double> &);
template bool skinny_fits<float>(const Gemm<float> &);
template bool skinny_fills<double>(const Gemm<double> &, int);
template bool skinny_fills<float>(const Gemm<float> &, int);
template int skinny_pick_splitk<double>(const Gemm<double> &, int);
template int skinny_pick_splitk<float>(const Gemm<float> &, int);
template hipError_t launch_skinny<double>(const Gemm<double> &, int, double *, hipStream_t);
template hipError_t launch_skinny<float>(const Gemm<float> &, int, float *, hipStream_t);

}  // namespace blasdsp
