// tile_body.h — one C tile of the MFMA tile GEMM, shared by the per-call
// kernel (gemm_tile.hip) and the grouped persistent kernel (gemm_grouped.hip).
//
// A workgroup of 4 or 8 waves computes one BM x BM tile of C = alpha op(A)
// op(B) + beta C with v_mfma_f64_16x16x4_f64 (FP64) or v_mfma_f32_16x16x4_f32
// (FP32):
//   * the waves form a 2 x 2 (or, with 8 waves, 4 x 2; 2 x 4 for TT) grid,
//     each owning a sub-tile of MFMA 16x16 accumulators;
//   * op(A) / op(B) tiles of BK = 128 bytes of K are staged global -> LDS with
//     16-byte vector loads along whichever dimension is contiguous in HBM, and
//     the LDS copy keeps that orientation; the MFMA fragment reads adapt to it,
//     so the transpose modes need no transpose pass and no transposing writes;
//   * LDS rows are padded: m/n-contiguous rows by 16 elements, which makes
//     their fragment reads bank-conflict-free, and k-contiguous rows by 16
//     bytes, which leaves a 2-way conflict (rows r and r + 8 of a 16-lane
//     read share banks; rocprofv3 SQ_LDS_BANK_CONFLICT on MI355X: zero for
//     two m/n-contiguous operands, one unit per K-contiguous operand).  An
//     XOR-swizzled k-contiguous layout measured zero conflicts but ran slower
//     (FP32 8192^3 -8 to -12 %, FP64 4096^3 -1 to -2 %): with four waves per
//     SIMD the conflicts' latency is hidden, and the swizzle's per-lane
//     addresses are not;
//   * two LDS stages: global loads for K tile k+1 are in flight while the
//     MFMAs consume tile k; one barrier per K tile;
//   * epilogue C = alpha * acc + beta * C fused into the store (C is not read
//     when beta == 0; for a short K range C's lines are prefetched at the
//     tile's start);
//   * operand loads are buffer loads at any element alignment of the leading
//     dimension; out-of-range rows and the ragged edge read as zero through
//     the buffer's range check (see Stager::load);
//   * interior tiles with K a multiple of BK take a specialised body with no
//     bounds checks (5 % faster at FP64 4096^3).
#pragma once

#include "internal.h"

namespace blasdsp {
namespace tile {

template <typename T>
struct Mma;

template <>
struct Mma<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    typedef double vec_t __attribute__((ext_vector_type(2)));
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // Output of the f64 16x16 block: lane group g = lane / 16 and register r
    // hold row g + 4 r, column lane % 16.
    static __device__ __forceinline__ int out_row(int g, int r) { return g + 4 * r; }
};

template <>
struct Mma<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    typedef float vec_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // f32 16x16 block: lane group g, register r hold row 4 g + r.
    static __device__ __forceinline__ int out_row(int g, int r) { return 4 * g + r; }
};

// Tile geometry.  NW waves (4 or 8) in a WGM x WGN grid (WGN 2, or 4 for
// TT with 8 waves); each owns a WM x WN sub-tile of MI x NJ MFMA 16x16
// accumulators.
template <typename T, int BM_, int NW_ = 4, int WGN_ = 2>
struct Cfg {
    static constexpr int BM = BM_;
    static constexpr int NW = NW_;
    static_assert(NW == 4 || NW == 8, "4 or 8 waves");
    static_assert(NW % WGN_ == 0 && (WGN_ == 2 || WGN_ == 4), "wave grid");
    static constexpr int BK = 128 / (int)sizeof(T);  // one 128-byte row of K
    static constexpr int V = 16 / (int)sizeof(T);    // elements per 16-byte vector
    static constexpr int SK = BK + V;                // k-contiguous LDS row stride
    static constexpr int SMN = BM + 16;              // m/n-contiguous LDS row stride
    static constexpr int BUF = (BM * SK > BK * SMN) ? BM * SK : BK * SMN;  // per operand stage
    static constexpr int THREADS = 64 * NW;
    static constexpr int WGM = NW / WGN_, WGN = WGN_;  // wave grid
    static constexpr int WM = BM / WGM, WN = BM / WGN;
    static constexpr int MI = WM / 16, NJ = WN / 16;
    static constexpr int LDS_ELEMS = 2 * 2 * BUF;    // [stage][A, B]
    // Waves per SIMD the register allocation must allow: the workgroups per CU
    // that the LDS admits (two of 72 KiB for BM = 128, four for BM = 64) times
    // NW / 4.  For FP64 4096^3 two 4-wave 128-tile workgroups per CU measured
    // 60 vs 44 TFLOP/s against one, and 8 waves (32 x 64 each, four per SIMD)
    // 68 vs 64 against 4 waves in the same specialised body.
    static constexpr int WPE = (BM == 128 ? 2 : 4) * NW / 4;
};

// Global -> register -> LDS staging of one operand tile.  KC: the stored tile
// is k-contiguous (BM rows of BK elements); otherwise m/n-contiguous (BK rows
// of BM elements).
template <typename Cf, typename T, bool KC>
struct Stager {
    using vec_t = typename Mma<T>::vec_t;
    static constexpr int BM = Cf::BM;
    static constexpr int TR = KC ? BM : Cf::BK;
    static constexpr int TC = KC ? Cf::BK : BM;
    static constexpr int CPR = TC / Cf::V;
    static constexpr int NCH = TR * CPR / Cf::THREADS;
    static constexpr int S = KC ? Cf::SK : Cf::SMN;
    static_assert(TR * CPR % Cf::THREADS == 0, "whole vectors per thread");

    vec_t r[NCH];
    int sh = 0;  // the loaded vectors start sh elements early (see load)

    // One element of a buffer load; an offset at or past the resource's
    // extent (0x7fffffff bytes) reads as zero without accessing memory.
    static __device__ __forceinline__ T load_elem(__amdgpu_buffer_rsrc_t rs, uint32_t off)
    {
        if constexpr (sizeof(T) == 8)
            return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0));
        else
            return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0));
    }

    // X: stored matrix with R rows, Cn columns, leading dimension ld; the tile
    // starts at stored (r0, c0), r0 < R and c0 < Cn (or Cn == 0: nothing is
    // read).  Elements outside the matrix read as zero.  Thread tid's vector
    // i is row tid / CPR + i * RSTEP, columns (tid % CPR) * V + [0, V).
    //
    // PTR = false (the leading dimension within full_offsets_fit, Cn >= V):
    // buffer loads from each vector row's workgroup-uniform origin (a scalar
    // resource) plus one 32-bit lane offset, one 16-byte load per vector at
    // any element alignment of ld (the HSA code object model enables
    // unaligned global access; the compiler itself lowers align-4 16-byte
    // loads to the same instruction).  A vector whose row is past R or whose
    // first column is past Cn gets an offset past the resource's extent,
    // which reads as zero without accessing memory.  A vector that starts
    // inside and ends past Cn (the K tail of a K-contiguous operand, the M or
    // N edge of an M/N-contiguous one) is loaded shifted back to end at
    // column Cn - 1, and store() shifts it into place with zeros behind.  The
    // code has no branches, so the loads of both operands stay in flight
    // together (a branch between them made the compiler wait for the first
    // operand's loads before issuing the second's).  FULL: the caller
    // guarantees the whole tile is inside.
    // PTR = true: 64-bit per-lane addresses, for leading dimensions whose row
    // offsets do not fit 32 bits and for matrices of fewer than V columns.
    template <bool FULL, bool PTR = false>
    __device__ __forceinline__ void load(const T *__restrict__ X, int64_t ld, int64_t R, int64_t Cn,
                                         int64_t r0, int64_t c0, int tid)
    {
        constexpr int RSTEP = Cf::THREADS / CPR;
        static_assert(Cf::THREADS % CPR == 0, "whole rows per step");
        const int row = tid / CPR, col = (tid % CPR) * Cf::V;
        if constexpr (!PTR) {
            constexpr uint32_t kOut = 0x80000000u;
            const int64_t gc = c0 + col;
            const bool col_in = FULL || gc < Cn;
            const int64_t over = gc + Cf::V - Cn;  // elements past the last column
            sh = (FULL || !col_in || over <= 0) ? 0 : (int)over;
            // The resource base sits V elements before the vector row's origin
            // so that a shifted vector's offset (at least V - sh > 0) stays
            // positive; the addresses read are never before the origin's row.
            const uint32_t off = (uint32_t)(((int64_t)row * ld + col + Cf::V - sh) * (int64_t)sizeof(T));
#pragma unroll
            for (int i = 0; i < NCH; ++i) {
                const uintptr_t base =
                    (uintptr_t)X +
                    (uintptr_t)(((r0 + (int64_t)i * RSTEP) * ld + c0 - Cf::V) * (int64_t)sizeof(T));
                const __amdgpu_buffer_rsrc_t rs =
                    __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, 0x7fffffff, 0x00020000);
                const bool live = FULL || (col_in && r0 + row + (int64_t)i * RSTEP < R);
                r[i] = __builtin_bit_cast(vec_t, __builtin_amdgcn_raw_buffer_load_b128(rs, live ? off : kOut, 0, 0));
            }
            return;
        }
        const bool cols_in = c0 + TC <= Cn;
        sh = 0;
#pragma unroll
        for (int i = 0; i < NCH; ++i) {
            const int64_t gr = r0 + row + (int64_t)i * RSTEP, gc = c0 + col;
            if (cols_in && gr < R) {
                __builtin_memcpy(&r[i], X + gr * ld + gc, sizeof(vec_t));  // 16 bytes at element alignment
            } else {
#pragma unroll
                for (int e = 0; e < Cf::V; ++e) r[i][e] = (gr < R && gc + e < Cn) ? X[gr * ld + gc + e] : T(0);
            }
        }
    }

    // Vector v loaded sh elements early: element e is v[e + sh], zero past
    // the end.
    static __device__ __forceinline__ vec_t unshift(vec_t v, int sh)
    {
        vec_t o;
#pragma unroll
        for (int e = 0; e < Cf::V; ++e) {
            T x = v[e];
#pragma unroll
            for (int t = 1; t < Cf::V; ++t) x = (sh == t) ? (e + t < Cf::V ? v[e + t] : T(0)) : x;
            o[e] = x;
        }
        return o;
    }

    // SHIFT: undo the load's shift of vectors that end past the last column.
    template <bool SHIFT>
    __device__ __forceinline__ void store(T *s, int tid) const
    {
#pragma unroll
        for (int i = 0; i < NCH; ++i) {
            const int c = tid + i * Cf::THREADS;
            const int row = c / CPR, col = (c % CPR) * Cf::V;
            *reinterpret_cast<vec_t *>(s + row * S + col) = SHIFT ? unshift(r[i], sh) : r[i];
        }
    }
};

// The buffer loads' lane offsets (up to BM rows of a stored operand tile)
// stay below 2^31 bytes for leading dimension ld.
template <typename T, int BM>
__host__ __device__ inline bool full_offsets_fit(int64_t ld)
{
    return ld <= (int64_t)((2147483647 - 4096) / ((int64_t)(BM - 1) * (int64_t)sizeof(T)));
}

// The tile body may take buffer loads (PTR = false) for g: both leading
// dimensions within full_offsets_fit and both stored operands at least one
// 16-byte vector wide (a vector that ends past the last column is loaded
// shifted back inside the row).
template <typename T, int BM>
inline bool buffer_ok(const Gemm<T> &g)
{
    constexpr int64_t V = 16 / (int64_t)sizeof(T);
    const int64_t ac = g.ta ? g.M : g.K, bc = g.tb ? g.K : g.N;
    return full_offsets_fit<T, BM>(g.lda) && full_offsets_fit<T, BM>(g.ldb) && ac >= V && bc >= V;
}

// True when tile (m0, n0) over K range [kb, ke) can take the FULL body: the
// tile inside C and the range a whole number of K tiles.
template <typename Cf, typename T>
__device__ __forceinline__ bool full_tile(const Gemm<T> &g, int64_t m0, int64_t n0, int64_t kb, int64_t ke)
{
    return m0 + Cf::BM <= g.M && n0 + Cf::BM <= g.N && (ke - kb) % Cf::BK == 0;
}

// Wave issue priority (s_setprio 1) over one section of each K tile, per
// precision and transpose mode, where it measured faster on MI355X with the
// 8-wave 128 tile (interleaved A/B on four boxes, bit-identical results):
//   over the MFMA section (mfma_prio): FP64 TN +4.5-5 %, FP32 TN +2.6-2.8 %,
//     TT +1.2 %; for FP64 NN/NT/TT and FP32 NN/NT it cost 1-3 %;
//   over the LDS stores and the barrier (store_prio): FP64 NN +1.0 %, NT
//     +1.8-5.5 %, TT +1.2 %; for FP64 TN and FP32 NT/TN it cost 3-5 %.
// Raising it for the whole K tile, or over the global loads, changed nothing.
template <typename T, int BM, int NW, bool TA, bool TB>
constexpr bool mfma_prio()
{
    return BM == 128 && NW == 8 && (sizeof(T) == 8 ? (TA && !TB) : TA);
}
template <typename T, int BM, int NW, bool TA, bool TB>
constexpr bool store_prio()
{
    return BM == 128 && NW == 8 && sizeof(T) == 8 && !(TA && !TB);
}

// Longest K range of a tile whose C lines are prefetched at the tile's start
// (see tile_body).
template <typename T>
constexpr int64_t kPrefetchCK = 2048;

// One BM x BM tile of C at (m0, n0) with NW waves, over K range [kb, ke)
// (kb a multiple of BK).  lds: Cfg::LDS_ELEMS elements.  Every wave of the
// workgroup must call it (it contains barriers).  FULL (see full_tile): no
// bounds checks anywhere.  PTR: 64-bit operand addressing (otherwise both
// leading dimensions must pass full_offsets_fit and both stored operands have
// at least V columns: buffer_ok).  W == nullptr: C = alpha
// acc + beta C; otherwise the raw sums go to W (M x N, leading dimension N), a
// split-K partial.  The MFMA sequence of every C element is the same for any
// NW, FULL and PTR, so those variants give bit-identical results.  tid: the
// thread's index in the workgroup (threadIdx.x when negative).
// Wave grid of the 8-wave tiles: 4 x 2 (each wave 32 rows x 64 columns, so
// it reads 2 op(A) and 4 op(B) fragment tiles per k step), except TT, which
// takes 2 x 4.  A K-contiguous operand's LDS rows have 2-way bank conflicts
// (op(A) in NN and NT, op(B) in NT and TT); TT's conflicting operand is then
// read half as often.  Measured on MI355X (bit-identical): DGEMM 4096^3 TT
// 67.8 -> 69.6 TFLOP/s, FP32 8192^3 TT +0.7 %; NN with 2 x 4 reads its
// conflicting op(A) twice as often, and NT has both operands K-contiguous.
constexpr bool kFlipTT = true;

template <typename T, int BM, bool TA, bool TB, int NW = 4, bool FULL = false, bool PTR = false>
__device__ __forceinline__ void tile_body(const Gemm<T> &g, int64_t m0, int64_t n0, T *lds, int64_t kb = 0,
                                          int64_t ke = -1, T *W = nullptr, int tid = -1)
{
    if (ke < 0) ke = g.K;
    if (tid < 0) tid = threadIdx.x;
    using Cf = Cfg<T, BM, NW, (kFlipTT && NW == 8 && TA && TB) ? 4 : 2>;
    using acc_t = typename Mma<T>::acc_t;
    constexpr int BK = Cf::BK, SK = Cf::SK, SMN = Cf::SMN, MI = Cf::MI, NJ = Cf::NJ, BUF = Cf::BUF;

    const int lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int wm = (wave / Cf::WGN) * Cf::WM, wn = (wave % Cf::WGN) * Cf::WN;

    // Stored geometry: A is K x M when transposed, else M x K; B is N x K when
    // transposed, else K x N.
    const int64_t AR = TA ? g.K : g.M, AC = TA ? g.M : g.K;
    const int64_t BR = TB ? g.N : g.K, BC = TB ? g.K : g.N;

    Stager<Cf, T, !TA> sa;
    Stager<Cf, T, TB> sb;
    auto load_tiles = [&](int64_t k0) {
        if (TA) sa.template load<FULL, PTR>(g.A, g.lda, AR, AC, k0, m0, tid);
        else    sa.template load<FULL, PTR>(g.A, g.lda, AR, AC, m0, k0, tid);
        if (TB) sb.template load<FULL, PTR>(g.B, g.ldb, BR, BC, n0, k0, tid);
        else    sb.template load<FULL, PTR>(g.B, g.ldb, BR, BC, k0, n0, tid);
    };
    auto stage = [&](int st, int op) { return lds + (st * 2 + op) * BUF; };

    acc_t acc[MI][NJ];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    const int nk = (int)(FULL ? (ke - kb) / BK : (ke - kb + BK - 1) / BK);
    constexpr bool SHIFT = !FULL && !PTR;
    // C lines for a short K range: one 4-byte load per 128-byte line of the
    // tile's C rows, issued with the first K tile's loads, brings C into the
    // caches while the K loop runs, so the epilogue's read of C does not wait
    // on HBM.  Lines outside C read nothing (buffer range check).  Measured
    // on MI355X on the 512-request serving-mix batch, whose tiles are mostly
    // short in K: reading C at all cost 12 % of the FP64 128 list's time and
    // 6 % of the FP32 128 list's (beta = 0 against beta = 1).  At FP64
    // 4096^3 the same prefetch was 0.4-2.7 % slower (a long K walk evicts the
    // lines again), hence the K bound.
    constexpr int kLinesPerRow = BM * (int)sizeof(T) / 128;
    constexpr int kPfPerThread = (BM * kLinesPerRow + Cf::THREADS - 1) / Cf::THREADS;
    uint32_t pf[kPfPerThread];
    const bool prefetch_c = !PTR && W == nullptr && g.beta != T(0) && ke - kb <= kPrefetchCK<T> &&
                            full_offsets_fit<T, BM>(g.ldc);
    if (prefetch_c) {
        const __amdgpu_buffer_rsrc_t rc =
            __builtin_amdgcn_make_buffer_rsrc((void *)(g.C + m0 * g.ldc + n0), (short)0, 0x7fffffff, 0x00020000);
#pragma unroll
        for (int q = 0; q < kPfPerThread; ++q) {
            const int li = tid + q * Cf::THREADS;
            const int row = li / kLinesPerRow, col = (li % kLinesPerRow) * (128 / (int)sizeof(T));
            const bool in = li < BM * kLinesPerRow && m0 + row < g.M && n0 + col < g.N;
            const uint32_t off = in ? (uint32_t)(((int64_t)row * g.ldc + col) * (int64_t)sizeof(T)) : 0x80000000u;
            pf[q] = __builtin_amdgcn_raw_buffer_load_b32(rc, off, 0, 0);
        }
    }
    load_tiles(kb);
    sa.template store<SHIFT>(stage(0, 0), tid);
    sb.template store<SHIFT>(stage(0, 1), tid);
    __syncthreads();
    // the prefetch loads were issued before the first K tile's, which have
    // been waited for: consuming the values here costs no wait
    if (prefetch_c) {
#pragma unroll
        for (int q = 0; q < kPfPerThread; ++q) asm volatile("" ::"v"(pf[q]));
    }

    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) load_tiles(kb + (int64_t)(kt + 1) * BK);

        const T *As = stage(cur, 0);
        const T *Bs = stage(cur, 1);
        if constexpr (mfma_prio<T, BM, NW, TA, TB>()) __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (int ks = 0; ks < BK / 4; ++ks) {
            const int k = ks * 4 + lk;
            T a[MI], b[NJ];
#pragma unroll
            for (int i = 0; i < MI; ++i) {
                const int m = wm + i * 16 + lr;
                a[i] = TA ? As[k * SMN + m] : As[m * SK + k];
            }
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                const int n = wn + j * 16 + lr;
                b[j] = TB ? Bs[n * SK + k] : Bs[k * SMN + n];
            }
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int j = 0; j < NJ; ++j) acc[i][j] = Mma<T>::mma(a[i], b[j], acc[i][j]);
        }

        if constexpr (mfma_prio<T, BM, NW, TA, TB>()) __builtin_amdgcn_s_setprio(0);
        if constexpr (store_prio<T, BM, NW, TA, TB>()) __builtin_amdgcn_s_setprio(1);
        if (more) {
            sa.template store<SHIFT>(stage(cur ^ 1, 0), tid);
            sb.template store<SHIFT>(stage(cur ^ 1, 1), tid);
        }
        __syncthreads();
        if constexpr (store_prio<T, BM, NW, TA, TB>()) __builtin_amdgcn_s_setprio(0);
    }

    const T alpha = g.alpha, beta = g.beta;
#pragma unroll
    for (int i = 0; i < MI; ++i) {
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int64_t col = n0 + wn + j * 16 + lr;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = m0 + wm + i * 16 + Mma<T>::out_row(lk, r);
                if (FULL || (row < g.M && col < g.N)) {
                    if (W) {
                        W[row * g.N + col] = acc[i][j][r];
                    } else {
                        T *p = g.C + row * g.ldc + col;
                        T v = alpha * acc[i][j][r];
                        if (beta != T(0)) v += beta * *p;
                        *p = v;
                    }
                }
            }
        }
    }
}

}  // namespace tile
}  // namespace blasdsp
