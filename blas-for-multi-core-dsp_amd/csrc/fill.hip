// fill.hip — device-side synthetic input generator.
//
// Counter-based: element (i, j) of a logical rows x cols matrix takes draw
// number i * cols + j of a splitmix64 stream keyed by the seed, so the same
// logical matrix results for any leading dimension or storage order, and any
// element can be produced independently (one thread per element).  Uniform
// and the integer patterns use only integer arithmetic and exact double
// operations, so they match a host implementation of the same definition bit
// for bit; normal (Box-Muller) and log-uniform use device log/cos/exp2.
#include "internal.h"

namespace blasdsp {
namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ULL;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBULL;
    x ^= x >> 31;
    return x;
}

__device__ __forceinline__ uint64_t stream_bits(uint64_t key, uint64_t c)
{
    return mix64(key + (c + 1) * 0x9E3779B97F4A7C15ULL);
}

__device__ __forceinline__ double unit01(uint64_t bits) { return (double)(bits >> 11) * 0x1.0p-53; }

__device__ double draw(int dist, uint64_t key, uint64_t idx)
{
    switch (dist) {
    case 0:
        return 2.0 * unit01(stream_bits(key, idx)) - 1.0;
    case 1: {
        const double u1 = 1.0 - unit01(stream_bits(key, 2 * idx));
        const double u2 = unit01(stream_bits(key, 2 * idx + 1));
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
    }
    case 2: {
        const uint64_t b0 = stream_bits(key, 2 * idx);
        const double u = unit01(stream_bits(key, 2 * idx + 1));
        const double mag = exp2(16.0 * u - 8.0);
        return (b0 >> 63) ? -mag : mag;
    }
    case 3:
        return (double)(idx % 11);
    case 4:
        return (double)(idx % 7);
    case 5:
        return 1.0;
    default:
        return 0.0;
    }
}

template <typename T>
__global__ void fill_kernel(T *X, int64_t rows, int64_t cols, int64_t ld, int trans, int dist,
                            uint64_t key, T pad)
{
    const int64_t srows = trans ? cols : rows;
    const int64_t scols = trans ? rows : cols;
    const int64_t total = srows * ld;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e / ld, c = e % ld;
        T v = pad;
        if (c < scols) {
            const int64_t i = trans ? c : r, j = trans ? r : c;
            v = (T)draw(dist, key, (uint64_t)(i * cols + j));
        }
        X[e] = v;
    }
}

}  // namespace

template <typename T>
hipError_t launch_fill(T *X, int64_t rows, int64_t cols, int64_t ld, int trans, int dist,
                       uint64_t seed, T pad, hipStream_t s)
{
    const int64_t srows = trans ? cols : rows;
    const int64_t total = srows * ld;
    if (total <= 0) return hipSuccess;
    // Same key derivation as the host generator.
    uint64_t key = seed * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL;
    key ^= key >> 30;
    key *= 0xBF58476D1CE4E5B9ULL;
    key ^= key >> 27;
    key *= 0x94D049BB133111EBULL;
    key ^= key >> 31;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL((fill_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, s, X, rows, cols, ld,
                       trans, dist, key, pad);
    return hipGetLastError();
}

template hipError_t launch_fill<double>(double *, int64_t, int64_t, int64_t, int, int, uint64_t, double,
                                        hipStream_t);
template hipError_t launch_fill<float>(float *, int64_t, int64_t, int64_t, int, int, uint64_t, float,
                                       hipStream_t);

}  // namespace blasdsp
