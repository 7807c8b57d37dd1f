// Device-side helpers shared by the HIP kernels of libdpca (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dpca {

// Elementwise kernels take a grid-stride loop: at most kEwBlocks workgroups
// of 256 threads, so the dispatch's work-item count stays far below the 2^32
// an AMD dispatch packet can describe (a 25 GB matrix has 2.5e10 bytes).
constexpr int64_t kEwBlocks = 1 << 20;
inline unsigned ew_blocks(int64_t n)
{
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < kEwBlocks ? b : kEwBlocks);
}
#define DPCA_EW_LOOP(idx, n)                                                          \
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < (n);    \
         idx += (int64_t)gridDim.x * blockDim.x)

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Tile (x, y) of a gx x gy tiling run as a 1-D grid of gx * gy workgroups.
// The dispatcher hands workgroup L to XCD L % 8; the tiles are ordered so
// that each XCD works through consecutive rows y (a product's operand for
// row y -- a sample segment of Q slices, an SNP split of U/W slices -- is
// then fetched into that XCD's L2 once rather than into all eight).  The
// last T mod 8 workgroups keep their own index.
__device__ __forceinline__ void xcd_tile(int gx, int &x, int &y)
{
    const int T = (int)gridDim.x, L = (int)blockIdx.x;
    const int T8 = T / 8 * 8;
    const int V = L < T8 ? (L % 8) * (T8 / 8) + L / 8 : L;
    x = V % gx;
    y = V / gx;
}

// 64-wide wavefront sum.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ long long wave_sum_ll(long long v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// 2-bit code of sample t (0..63) within 16 packed bytes held as 4 dwords.
__device__ __forceinline__ uint32_t code_at(const uint32_t (&w)[4], int t)
{
    return (w[t >> 4] >> (2 * (t & 15))) & 3u;
}

// Select one of four doubles by a 2-bit code without indexing registers.
__device__ __forceinline__ double sel4(uint32_t c, double v0, double v1, double v2,
                                       double v3)
{
    double lo = (c & 1u) ? v1 : v0;
    double hi = (c & 1u) ? v3 : v2;
    return (c & 2u) ? hi : lo;
}

}  // namespace dpca
