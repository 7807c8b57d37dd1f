// Mixed-precision form of one pass  A A^T Q = Z^T (Z Q)  on int8 MFMA
// (v_mfma_i32_16x16x64_i8).  The genotype codes enter the products as exact
// int8 values -- h, the allele count (0 when missing), and m, the
// non-missing indicator -- with Z_si = d_s (h_si - mu_s m_si); the FP64
// right-hand operand is split into three int8 fixed-point slices with one
// power-of-two scale per column and group of rows, so every product sums
// exactly in int32 and only the final combination rounds.
//
//   T = Z Q      two int8 rows per SNP (h and m) against three slices of Q
//                (scale per sample segment); T_s = d_s (H_s - mu_s M_s) in
//                FP64 per segment, FP32 partials, segments summed in FP64.
//   Z^T T        two int8 K slots per SNP (h and m) against three slices of
//                U = d T and W = -d mu T (scale per SNP split); splits
//                summed in FP64.
#include "dpca_internal.h"
#include "k_common.h"

#include <mutex>

namespace dpca {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));


// ----------------------------------------------------- T segments -> T -----
// T[s][c] = sum of the sample-segment partials of Z Q, in FP64.
// T = the sum of the segment partials (FP64, segment order); with smax,
// also smax[split * kp + c] = max over the split's rows of f_s |T_sc| as
// k_split_max (below) forms it, in the same pass over T.
// Block: 4 x kp threads over rb rows (rb divides 128, so a block stays
// inside one split: rps is a multiple of 128).
__global__ void k_tprep(const float *__restrict__ Tpart, int nseg, int64_t rows, int kp, int rb,
                        const double *__restrict__ mu, const double *__restrict__ dsc,
                        int64_t rps, unsigned *__restrict__ smax, double *__restrict__ T)
{
    __shared__ float red[256];
    const int tid = threadIdx.x, rl = tid / kp, c = tid % kp;
    const int64_t r0 = (int64_t)blockIdx.x * rb;
    double m = 0;
#pragma unroll 4
    for (int64_t s = r0 + rl; s < r0 + rb; s += 4) {
        const int64_t idx = s * kp + c;
        double t = 0;
        for (int p = 0; p < nseg; p++) t += (double)Tpart[(int64_t)p * rows * kp + idx];
        T[idx] = t;
        if (smax) {
            const double f = fabs(dsc[s]) * fmax(1.0, fabs(mu[s]));
            m = fmax(m, f * fabs(t));
        }
    }
    if (!smax) return;  // the same for the whole block
    float f = (float)m;
    if ((double)f < m) f = __uint_as_float(__float_as_uint(f) + 1u);  // round up
    red[tid] = f;
    __syncthreads();
    if (rl == 0) {
        for (int j = 1; j < 4; j++) f = fmaxf(f, red[j * kp + c]);
        if (f > 0) atomicMax(&smax[(r0 / rps) * kp + c], __float_as_uint(f));
    }
}

hipError_t launch_tprep(hipStream_t st, const float *Tpart, int nseg, int64_t rows, int kp,
                        double *T, const double *mu, const double *dsc, int64_t rps,
                        unsigned *smax)
{
    if (rows % 128 || kp * 4 > 256 || (smax && (rps % 128 || rps < 128)))
        return hipErrorInvalidValue;
    if (smax) {
        // a fill kernel, not a memset node: safe inside a captured graph
        const int64_t nsplit = (rows + rps - 1) / rps;
        hipError_t e = launch_fill_u8(st, reinterpret_cast<uint8_t *>(smax), nsplit * kp * 4, 0);
        if (e != hipSuccess) return e;
    }
    // short blocks of rows: >= ~2048 workgroups where the rows allow
    int rb = 128;
    while (rb > 16 && rows / rb < 2048) rb /= 2;
    k_tprep<<<(unsigned)(rows / rb), 4 * kp, 0, st>>>(Tpart, nseg, rows, kp, rb, mu, dsc, rps, smax,
                                                       T);
    return hipGetLastError();
}

hipError_t launch_reduce_f32(hipStream_t st, const float *part, int splits, int64_t n,
                             double *out, int accumulate)
{
    return launch_sum_parts_f32(st, part, splits, n, out, accumulate);
}

// ============================================== int8 slices and their tiles ==
// An FP64 operand x (Q, or U = d T and W = -d mu T) enters the MFMA as three
// int8 slices of a fixed-point form with one power-of-two scale s per column
// and group of rows (one workgroup's accumulation range):
// x = s (q0 + q1 2^-7 + q2 2^-14) + e, |e| <= s 2^-15, s >= max|x| / 127.
// The slice products sum exactly in int32; the slices are combined in FP64
// when the group ends.
//
// Column tiles.  The three slices of the k right-hand columns are laid out
// as 16-wide MFMA B tiles ("virtual tiles", VT = 3F + RT of them): columns
// 16c..16c+15, c < F, take three tiles each (tile 3c + j holds slice j in
// lane c%16); the R = k - 16F columns left over (R <= 10) share RT =
// ceil(3R / 16) tiles: virtual column w = jR + x holds slice j of column
// 16F + x, in tile 3F + w/16, lane w%16.  So k = 40 takes 8 tiles, not 9,
// and k = 20 four, not six.  k with R > 10 is padded to whole tiles.
//
// MFMA operand layout (v_mfma_i32_16x16x64_i8, checked on the device): lane
// l = 16g + r holds A[r][16g..16g+15] and B[16g..16g+15][r] as bytes and
// receives D[4g + q][r].  Which sample (or SNP) occupies which K byte only
// has to agree between A and B.

typedef int i32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ i32x4 mfma_i8(const u32x4 &a, const u32x4 &b, const i32x4 &c)
{
    return __builtin_amdgcn_mfma_i32_16x16x64_i8(__builtin_bit_cast(i32x4, a),
                                                 __builtin_bit_cast(i32x4, b), c, 0, 0, 0);
}

// power-of-two scale s with max / s < 127 (s = 1 for an all-zero column)
__device__ __forceinline__ double slice_scale(float mx)
{
    int e = 0;
    (void)frexp((double)mx / 127.0, &e);
    return ldexp(1.0, e);
}

// |x| <= 127 s: three int8 slices of x / s, each residual scaled by 2^7.
__device__ __forceinline__ void slices3(double x, double inv_s, int q[3])
{
    double v = x * inv_s;              // exact: inv_s is a power of two
    const double a = rint(v);
    v = (v - a) * 128.0;               // exact, |v| <= 64
    const double b = rint(v);
    v = (v - b) * 128.0;
    q[0] = (int)a;
    q[1] = (int)b;
    q[2] = (int)rint(v);
}

// Slice j and column of virtual column (tile v, lane r); column -1 = empty.
__device__ __forceinline__ void vcol(int v, int r, int F, int R, int &j, int &col)
{
    if (v < 3 * F) {
        j = v % 3;
        col = (v / 3) * 16 + r;
        return;
    }
    const int w = (v - 3 * F) * 16 + r;
    j = R ? w / R : 3;
    col = (j < 3) ? 16 * F + w % R : -1;
}

// Value of slice j of column 16F + x for this lane's D rows: the remainder
// tiles' accumulators are spread over lanes 16g + (jR + x) % 16.
template <int F, int RT>
__device__ __forceinline__ int rem_slice(const i32x4 (&acc)[3 * F + RT], int q, int j, int x,
                                         int R, int g)
{
    const int w = j * R + x;
    const int src = 16 * g + (w & 15);
    int val = 0;
#pragma unroll
    for (int u = 0; u < RT; u++) {
        const int y = __shfl(acc[3 * F + u][q], src, 64);
        if ((w >> 4) == u) val = y;
    }
    return val;
}

// Combined FP64 value sum_j 2^-7j slice_j of output column 16 oc + r for
// D row 4g + q (zero for padding columns).
template <int F, int RT>
__device__ __forceinline__ double combine(const i32x4 (&acc)[3 * F + RT], int oc, int q, int r,
                                          int R, int g)
{
    if (oc < F)
        return (double)acc[3 * oc][q] + ldexp((double)acc[3 * oc + 1][q], -7) +
               ldexp((double)acc[3 * oc + 2][q], -14);
    // remainder columns: every lane joins the shuffles, lanes r >= R drop out
    const int x = r < R ? r : 0;
    const int s0 = rem_slice<F, RT>(acc, q, 0, x, R, g);
    const int s1 = rem_slice<F, RT>(acc, q, 1, x, R, g);
    const int s2 = rem_slice<F, RT>(acc, q, 2, x, R, g);
    if (r >= R) return 0.0;
    return (double)s0 + ldexp((double)s1, -7) + ldexp((double)s2, -14);
}

// Same from slices 0 and 1 only (the slice-2 tiles of a product that needs
// 2^-7 less precision are not computed).
template <int F, int RT>
__device__ __forceinline__ double combine2(const i32x4 (&acc)[3 * F + RT], int oc, int q, int r,
                                           int R, int g)
{
    if (oc < F) return (double)acc[3 * oc][q] + ldexp((double)acc[3 * oc + 1][q], -7);
    const int x = r < R ? r : 0;
    const int s0 = rem_slice<F, RT>(acc, q, 0, x, R, g);
    const int s1 = rem_slice<F, RT>(acc, q, 1, x, R, g);
    if (r >= R) return 0.0;
    return (double)s0 + ldexp((double)s1, -7);
}

// smax[split * kp + c] = max over rows of the split of f_s |T_sc|, with
// f_s = |d_s| max(1, |mu_s|) (covers |U| = |d T| and |W| = |d mu T|), or
// f_s = 1 when dsc is null.
// Block: 4 x kp threads over 128 rows of one split; one atomic per column.
// smax must hold zeros (or smaller maxima) on entry.
__global__ void k_split_max(const double *__restrict__ T, int64_t rows, int kp,
                            const double *__restrict__ mu, const double *__restrict__ dsc,
                            int64_t rps, unsigned *__restrict__ smax)
{
    __shared__ float red[256];
    const int tid = threadIdx.x, rl = tid / kp, c = tid % kp;
    const int64_t split = blockIdx.y;
    const int64_t r0 = split * rps + (int64_t)blockIdx.x * 128;
    int64_t r1 = r0 + 128;
    if (r1 > (split + 1) * rps) r1 = (split + 1) * rps;
    if (r1 > rows) r1 = rows;
    double m = 0;
    for (int64_t s = r0 + rl; s < r1; s += 4) {
        const double f = dsc ? fabs(dsc[s]) * fmax(1.0, fabs(mu[s])) : 1.0;
        m = fmax(m, f * fabs(T[s * kp + c]));
    }
    float f = (float)m;
    if ((double)f < m) f = __uint_as_float(__float_as_uint(f) + 1u);  // round up
    red[tid] = f;
    __syncthreads();
    if (rl == 0) {
        for (int j = 1; j < 4; j++) f = fmaxf(f, red[j * kp + c]);
        if (f > 0) atomicMax(&smax[split * kp + c], __float_as_uint(f));
    }
}

// ====================================================== Z^T T on int8 MFMA ==
// Z^T T = sum_s [h_si U_s + m_si W_s]: the codes give h (0/1/2) and m (0/1)
// as exact int8 K slots; U and W share one scale per (SNP split, column) so
// both slots feed one accumulator (|sum| <= 381 per SNP).  K step t covers 32
// SNPs; lane group g takes SNPs 32t + 8g .. +7 with K bytes
// [h(0..3), h(4..7), m(0..3), m(4..7)].
//
// UW8 layout: per 8-SNP group G, per virtual tile v, per lane r one 16-byte
// word {U_j(8G..+3), U_j(8G+4..+7), W_j(8G..+3), W_j(8G+4..+7)} of the slice
// and column of (v, r), at UW8[(G * VT + v) * 16 + r].

__global__ void k_uw8(const double *__restrict__ T, int64_t rows, int kp, int F, int R, int VT,
                      const double *__restrict__ mu, const double *__restrict__ dsc, int64_t rps,
                      const unsigned *__restrict__ smax, u32x4 *__restrict__ UW8)
{
    DPCA_EW_LOOP(idx, rows / 8 * VT * 16) {
        const int64_t G = idx / (VT * 16);
        const int v = (int)(idx % (VT * 16)) / 16, r = (int)(idx % 16);
        int j, c;
        vcol(v, r, F, R, j, c);
        uint32_t w[4] = {0, 0, 0, 0};
        if (c >= 0) {
            const int64_t split = (G * 8) / rps;
            const double inv = 1.0 / slice_scale(__uint_as_float(smax[split * kp + c]));
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int64_t row = G * 8 + e;
                const double t = T[row * kp + c];
                int qu[3], qw[3];
                slices3(dsc[row] * t, inv, qu);
                slices3(-(dsc[row] * mu[row]) * t, inv, qw);
                const int sh = 8 * (e & 3);
                w[e >> 2] |= (uint32_t)(qu[j] & 0xFF) << sh;
                w[2 + (e >> 2)] |= (uint32_t)(qw[j] & 0xFF) << sh;
            }
        }
        UW8[idx] = u32x4{w[0], w[1], w[2], w[3]};
    }
}

hipError_t launch_uw8(hipStream_t st, const double *T, int64_t rows, int k, int kp,
                      const double *mu, const double *dsc, int64_t rps, unsigned *smax, void *UW8,
                      bool have_max)
{
    const ColTiles ct = col_tiles(k);
    if (rows % 128 || rps % 128 || rps < 128 || kp * 4 > 256 || ct.kp() != kp)
        return hipErrorInvalidValue;
    const int64_t nsplit = (rows + rps - 1) / rps;
    if (!have_max) {
        // a fill kernel, not a memset node: safe inside a captured graph
        hipError_t e = launch_fill_u8(st, reinterpret_cast<uint8_t *>(smax), nsplit * kp * 4, 0);
        if (e != hipSuccess) return e;
        dim3 g1((unsigned)((rps + 127) / 128), (unsigned)nsplit);
        k_split_max<<<g1, 4 * kp, 0, st>>>(T, rows, kp, mu, dsc, rps, smax);
    }
    const int64_t n = rows / 8 * ct.vt() * 16;
    k_uw8<<<ew_blocks(n), 256, 0, st>>>(T, rows, kp, ct.F, ct.R, ct.vt(), mu, dsc,
                                                        rps, smax, reinterpret_cast<u32x4 *>(UW8));
    return hipGetLastError();
}

// Workgroup: 8 waves x SA tiles x 16 samples (SA = 4: 512 samples = 128
// packed bytes per row) and one SNP split (blockIdx.y); 128-SNP chunks of
// packed bytes staged in LDS transposed to [byte column][4-SNP group], so a
// lane reads the codes of 4 consecutive SNPs for its sample with one dword
// load, and the chunk's UW8 words copied to LDS.
template <int SA, int F, int RT>
__global__ __launch_bounds__(512) void k_zt_i8(const uint8_t *__restrict__ packed, int64_t ldp,
                                               const u32x4 *__restrict__ UW8,
                                               const unsigned *__restrict__ smax, int64_t rows,
                                               int64_t rows_per_split, int64_t Mpad, int R,
                                               float *__restrict__ part)
{
    constexpr int VT = 3 * F + RT;           // virtual column tiles
    constexpr int CT = F + (RT > 0 ? 1 : 0); // output column tiles
    constexpr int KP = CT * 16;
    constexpr int TB = SA * 16 * 8 / 4;    // packed bytes per row in this tile
    constexpr int SC = 128;                // SNPs per chunk
    constexpr int NUWT = (SC / 8) * VT * 16;  // UW8 words per chunk
    // two buffers: chunk i is multiplied out of one while chunk i+1 lands in
    // the other (one barrier per chunk).  pk rows are byte columns, their
    // dwords 4-SNP groups, stored at column sg ^ (row / 4) in rows of 36
    // dwords: both the transposing stores (one group, many rows per
    // half-wave) and the paired reads (4 rows x 2 groups) then touch
    // distinct banks.  Groups 2x, 2x+1 stay one aligned 8-byte pair, in
    // swapped order when row / 4 is odd.
    constexpr int PKW = SC / 4 + 4;
    static_assert(SA % 2 == 0, "row / 4 parity must follow the tile index");
    __shared__ uint32_t pk[2][TB][PKW];
    __shared__ u32x4 uw[2][NUWT];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int r = lane & 15, g = lane >> 4;
    const int64_t mb = (int64_t)blockIdx.x * TB;
    const int64_t sbeg = (int64_t)blockIdx.y * rows_per_split;
    int64_t send = sbeg + rows_per_split;
    if (send > rows) send = rows;
    const uint32_t HT = 0x02010000u, MT = 0x01010001u;  // h, m by 2-bit code

    i32x4 acc[SA][VT];
#pragma unroll
    for (int a = 0; a < SA; a++)
#pragma unroll
        for (int v = 0; v < VT; v++) acc[a][v] = i32x4{0, 0, 0, 0};

    constexpr int NPK = (SC / 4) * (TB / 4) / 512;   // 4x4 byte blocks per thread
    constexpr int NUW = (NUWT + 511) / 512;          // UW8 words per thread
    static_assert((SC / 4) * (TB / 4) % 512 == 0, "tile");
    // software pipeline: the next chunk is loaded into registers while the
    // current one is multiplied, then stored into the other LDS buffer
    uint32_t rpk[NPK][4];
    u32x4 ruw[NUW];
#define DPCA_ZT8_LOADP(c0, dst)                                                             \
    do {                                                                                    \
        _Pragma("unroll") for (int u = 0; u < NPK; u++) {                                  \
            const int it_ = tid + 512 * u;                                                  \
            const int sg_ = it_ / (TB / 4), bq_ = it_ % (TB / 4);                           \
            const uint8_t *p_ = packed + ((c0) + 4 * sg_) * ldp + mb + 4 * bq_;             \
            _Pragma("unroll") for (int j = 0; j < 4; j++) dst[u][j] =                      \
                *reinterpret_cast<const uint32_t *>(p_ + j * ldp);                          \
        }                                                                                   \
    } while (0)
#define DPCA_ZT8_LOADU(c0)                                                                  \
    do {                                                                                    \
        const u32x4 *q_ = UW8 + ((c0) / 8) * VT * 16;                                       \
        _Pragma("unroll") for (int u = 0; u < NUW; u++) {                                  \
            const int it_ = tid + 512 * u;                                                  \
            if (it_ < NUWT) ruw[u] = q_[it_];                                               \
        }                                                                                   \
    } while (0)
    // 4x4 byte transpose of each block into [byte column][SNP group]
#define DPCA_ZT8_STORE(buf)                                                                 \
    do {                                                                                    \
        _Pragma("unroll") for (int u = 0; u < NPK; u++) {                                  \
            const int it_ = tid + 512 * u;                                                  \
            const int sg_ = it_ / (TB / 4), bq_ = it_ % (TB / 4);                           \
            const uint32_t r0_ = rpk[u][0], r1_ = rpk[u][1], r2_ = rpk[u][2], r3_ = rpk[u][3]; \
            const uint32_t lo01_ = __builtin_amdgcn_perm(r1_, r0_, 0x05010400u);            \
            const uint32_t hi01_ = __builtin_amdgcn_perm(r1_, r0_, 0x07030602u);            \
            const uint32_t lo23_ = __builtin_amdgcn_perm(r3_, r2_, 0x05010400u);            \
            const uint32_t hi23_ = __builtin_amdgcn_perm(r3_, r2_, 0x07030602u);            \
            const int col_ = sg_ ^ (bq_ & 31);                                              \
            pk[buf][4 * bq_ + 0][col_] = __builtin_amdgcn_perm(lo23_, lo01_, 0x05040100u);  \
            pk[buf][4 * bq_ + 1][col_] = __builtin_amdgcn_perm(lo23_, lo01_, 0x07060302u);  \
            pk[buf][4 * bq_ + 2][col_] = __builtin_amdgcn_perm(hi23_, hi01_, 0x05040100u);  \
            pk[buf][4 * bq_ + 3][col_] = __builtin_amdgcn_perm(hi23_, hi01_, 0x07060302u);  \
        }                                                                                   \
        _Pragma("unroll") for (int u = 0; u < NUW; u++) {                                  \
            const int it_ = tid + 512 * u;                                                  \
            if (it_ < NUWT) uw[buf][it_] = ruw[u];                                          \
        }                                                                                   \
    } while (0)
    {
        const int64_t c0 = sbeg < send ? sbeg : 0;
        DPCA_ZT8_LOADP(c0, rpk);
        DPCA_ZT8_LOADU(c0);
        DPCA_ZT8_STORE(0);
    }
    __syncthreads();

    int cur = 0;
    for (int64_t s0 = sbeg; s0 < send; s0 += SC, cur ^= 1) {
        {
            // unconditional (re-reads the last chunk at the end)
            const int64_t nx = (s0 + SC < send) ? s0 + SC : s0;
            DPCA_ZT8_LOADP(nx, rpk);
            DPCA_ZT8_LOADU(nx);
        }
#pragma unroll
        for (int t = 0; t < SC / 32; t++) {
            u32x4 av[SA];
#pragma unroll
            for (int a = 0; a < SA; a++) {
                const int i = wave * SA * 16 + a * 16 + r;  // sample within the tile
                const int sh = 2 * (i & 3);
                const int sw = (wave * SA + a) & 30;         // (i / 16) & ~1
                const uint2 w2 = *reinterpret_cast<const uint2 *>(
                    &pk[cur][i >> 2][(8 * t + 2 * g) ^ sw]);
                const uint32_t x0 = (a & 1) ? w2.y : w2.x;   // group 8t + 2g
                const uint32_t x1 = (a & 1) ? w2.x : w2.y;   // group 8t + 2g + 1
                const uint32_t c0 = (x0 >> sh) & 0x03030303u;
                const uint32_t c1 = (x1 >> sh) & 0x03030303u;
                av[a] = u32x4{__builtin_amdgcn_perm(0u, HT, c0), __builtin_amdgcn_perm(0u, HT, c1),
                              __builtin_amdgcn_perm(0u, MT, c0), __builtin_amdgcn_perm(0u, MT, c1)};
            }
#pragma unroll
            for (int v = 0; v < VT; v++) {
                const u32x4 b = uw[cur][((4 * t + g) * VT + v) * 16 + r];
#pragma unroll
                for (int a = 0; a < SA; a++) acc[a][v] = mfma_i8(av[a], b, acc[a][v]);
            }
        }
        DPCA_ZT8_STORE(cur ^ 1);  // the other buffer: last read before the previous barrier
        __syncthreads();
    }
#undef DPCA_ZT8_STORE
#undef DPCA_ZT8_LOADU
#undef DPCA_ZT8_LOADP
    float *out = part + ((int64_t)blockIdx.y * Mpad + 4 * mb + wave * SA * 16) * KP;
#pragma unroll
    for (int oc = 0; oc < CT; oc++) {
        const int col = oc * 16 + r;
        const double s = slice_scale(__uint_as_float(smax[blockIdx.y * KP + col]));
#pragma unroll
        for (int a = 0; a < SA; a++)
#pragma unroll
            for (int q = 0; q < 4; q++)
                out[(a * 16 + 4 * g + q) * KP + col] =
                    (float)(s * combine<F, RT>(acc[a], oc, q, r, R, g));
    }
}

// ========================================================= Z Q on int8 MFMA ==
// T = Z Q with T_s = d_s (sum_i h_si Q_i - mu_s sum_i m_si Q_i): the codes
// give two exact int8 rows per SNP (h and m) against the slices of Q (one
// scale per sample segment and column), so each segment's sums are exact in
// int32 (|sum| <= 254 * 127 per 64 samples); the slices, the two rows and
// the per-SNP factors are combined in FP64 at the end of the segment and
// written as an FP32 partial of T.
//
// Q8 layout: per 16-sample group G, per virtual tile v, per lane r one
// 16-byte word of the slice and column of (v, r), at Q8[(G * VT + v) * 16 + r];
// byte e holds sample 16G + 4 (e & 3) + (e >> 2), the order in which k_za_i8
// unpacks a packed dword (its dword q holds the bit pairs q of the four
// bytes); qmax[seg * kp + c] the column's max |Q| over the segment.
__global__ void k_q8(const double *__restrict__ Q, int64_t Mpad, int kp, int F, int R, int VT,
                     int64_t seg_samples, const unsigned *__restrict__ qmax,
                     u32x4 *__restrict__ Q8)
{
    DPCA_EW_LOOP(idx, Mpad / 16 * VT * 16) {
        const int64_t G = idx / (VT * 16);
        const int v = (int)(idx % (VT * 16)) / 16, r = (int)(idx % 16);
        int j, c;
        vcol(v, r, F, R, j, c);
        uint32_t w[4] = {0, 0, 0, 0};
        if (c >= 0) {
            const int64_t seg = (G * 16) / seg_samples;
            const double inv = 1.0 / slice_scale(__uint_as_float(qmax[seg * kp + c]));
#pragma unroll
            for (int e = 0; e < 16; e++) {
                int q[3];
                slices3(Q[(G * 16 + 4 * (e & 3) + (e >> 2)) * kp + c], inv, q);
                w[e >> 2] |= (uint32_t)(q[j] & 0xFF) << (8 * (e & 3));
            }
        }
        Q8[idx] = u32x4{w[0], w[1], w[2], w[3]};
    }
}

hipError_t launch_q8(hipStream_t st, const double *Q, int64_t Mpad, int k, int kp,
                     int64_t seg_bytes, unsigned *qmax, void *Q8)
{
    const ColTiles ct = col_tiles(k);
    const int64_t segs = 4 * seg_bytes;  // samples per segment
    if (Mpad % 256 || segs % 256 || kp * 4 > 256 || ct.kp() != kp) return hipErrorInvalidValue;
    const int64_t nseg = (Mpad + segs - 1) / segs;
    hipError_t e = launch_fill_u8(st, reinterpret_cast<uint8_t *>(qmax), nseg * kp * 4, 0);
    if (e != hipSuccess) return e;
    dim3 g1((unsigned)((segs + 127) / 128), (unsigned)nseg);
    k_split_max<<<g1, 4 * kp, 0, st>>>(Q, Mpad, kp, nullptr, nullptr, segs, qmax);
    const int64_t n = Mpad / 16 * ct.vt() * 16;
    k_q8<<<ew_blocks(n), 256, 0, st>>>(Q, Mpad, kp, ct.F, ct.R, ct.vt(), segs, qmax,
                                                       reinterpret_cast<u32x4 *>(Q8));
    return hipGetLastError();
}

// Workgroup: 8 waves x ST tiles x 16 SNPs, one sample segment (blockIdx.y).
// Per chunk of CB packed bytes per row (4 CB samples), lane group g owns the
// CB samples of its CB / 4 packed bytes; at K step t it supplies samples
// CB g + 16t .. +15.  The chunk's Q8 words are staged in LDS.
template <int ST, int F, int RT, bool SKR>
__global__ __launch_bounds__(512) void k_za_i8(const uint8_t *__restrict__ packed, int64_t ldp,
                                               const u32x4 *__restrict__ Q8,
                                               const unsigned *__restrict__ qmax,
                                               const double *__restrict__ mu,
                                               const double *__restrict__ dsc, int64_t seg_bytes,
                                               int64_t Npad, int R, float *__restrict__ Tpart)
{
    constexpr int VT = 3 * F + RT;
    constexpr int CT = F + (RT > 0 ? 1 : 0);
    constexpr int KP = CT * 16;
    // packed bytes per chunk: 128 where two LDS buffers of its Q8 words
    // (CB * VT * 64 bytes each) fit the 160 KB of LDS, else 64
    constexpr int CB = (VT <= 10) ? 128 : 64;
    constexpr int NW = CB / 64;                 // 16-byte words per lane and row
    constexpr int NQT = (CB / 4) * VT * 16;     // Q8 words per chunk
    constexpr int NQ = (NQT + 511) / 512;
    __shared__ u32x4 qs[2][NQT];               // double buffer, one barrier per chunk
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int r = lane & 15, g = lane >> 4;
    const int64_t base = (int64_t)blockIdx.x * (8 * ST * 16) + wave * (ST * 16);
    const int64_t b0 = (int64_t)blockIdx.y * seg_bytes;
    int64_t b1 = b0 + seg_bytes;
    if (b1 > ldp) b1 = ldp;
    const uint32_t MT = 0x01010001u;  // m by 2-bit code
    // Per-row A values 63 h - mh with mh = rint(63 mu) (exact int8, 0 when
    // missing), as a byte table by 2-bit code: T_s = d_s (sum (63 h - mh m) Q
    // / 63 - delta_s sum m Q), delta_s = mu_s - mh / 63, |delta_s| <= 1/126,
    // so the mask product needs only slices 0 and 1 (2^-7 less precision).
    uint32_t tab[ST];
#pragma unroll
    for (int s = 0; s < ST; s++) {
        const int mh = (int)rint(63.0 * mu[base + s * 16 + r]);  // mu in [0, 2]
        tab[s] = ((uint32_t)(-mh) & 0xFFu) | (((uint32_t)(63 - mh) & 0xFFu) << 16) |
                 (((uint32_t)(126 - mh) & 0xFFu) << 24);
    }
    // tiles holding only slice 2 (of whole columns, and the last remainder
    // tile when SKR) are left out of the mask product
    auto slice2_only = [](int v) constexpr {
        return v < 3 * F ? v % 3 == 2 : (SKR && v == 3 * F + RT - 1);
    };

    i32x4 acch[ST][VT], accm[ST][VT];
#pragma unroll
    for (int s = 0; s < ST; s++)
#pragma unroll
        for (int v = 0; v < VT; v++) {
            acch[s][v] = i32x4{0, 0, 0, 0};
            accm[s][v] = i32x4{0, 0, 0, 0};
        }

    // software pipeline: the next chunk's Q8 words and packed bytes are
    // loaded into registers while the current chunk is multiplied; the Q8
    // words then go to the other LDS buffer
    u32x4 rq[NQ], rp[ST][NW];
#define DPCA_ZA8_LOAD(c0)                                                                   \
    do {                                                                                    \
        const u32x4 *src_ = Q8 + ((c0) / 4) * VT * 16;                                      \
        _Pragma("unroll") for (int u = 0; u < NQ; u++) {                                   \
            const int it_ = tid + 512 * u;                                                  \
            if (it_ < NQT) rq[u] = src_[it_];                                               \
        }                                                                                   \
        _Pragma("unroll") for (int s = 0; s < ST; s++)                                     \
            _Pragma("unroll") for (int n = 0; n < NW; n++) rp[s][n] =                      \
                *reinterpret_cast<const u32x4 *>(packed + (base + s * 16 + r) * ldp + (c0) + \
                                                 (CB / 4) * g + 16 * n);                    \
    } while (0)
#define DPCA_ZA8_STORE(buf)                                                                 \
    do {                                                                                    \
        _Pragma("unroll") for (int u = 0; u < NQ; u++) {                                   \
            const int it_ = tid + 512 * u;                                                  \
            if (it_ < NQT) qs[buf][it_] = rq[u];                                            \
        }                                                                                   \
    } while (0)
    DPCA_ZA8_LOAD(b0 < b1 ? b0 : 0);
    DPCA_ZA8_STORE(0);
    // where the registers allow, the packed codes (HBM) run two chunks
    // ahead and the Q8 words (L2) one
    constexpr bool PF2 = ST * VT <= 16;
    u32x4 rp2[ST][NW];
    if constexpr (PF2) {
        const int64_t c2 = (b0 + CB < b1) ? b0 + CB : b0;
#pragma unroll
        for (int s = 0; s < ST; s++)
#pragma unroll
            for (int n = 0; n < NW; n++)
                rp2[s][n] = *reinterpret_cast<const u32x4 *>(packed + (base + s * 16 + r) * ldp + c2 +
                                                             (CB / 4) * g + 16 * n);
    }
    __syncthreads();

    int cur = 0;
    for (int64_t cb = b0; cb < b1; cb += CB, cur ^= 1) {
        u32x4 pk[ST][NW];
#pragma unroll
        for (int s = 0; s < ST; s++)
#pragma unroll
            for (int n = 0; n < NW; n++) pk[s][n] = rp[s][n];
        {
            // unconditional (re-reads the last chunk at the end)
            const int64_t nx = (cb + CB < b1) ? cb + CB : cb;
            if constexpr (PF2) {
                const int64_t n2 = (cb + 2 * CB < b1) ? cb + 2 * CB : nx;
                const u32x4 *src_ = Q8 + (nx / 4) * VT * 16;
#pragma unroll
                for (int u = 0; u < NQ; u++) {
                    const int it_ = tid + 512 * u;
                    if (it_ < NQT) rq[u] = src_[it_];
                }
#pragma unroll
                for (int s = 0; s < ST; s++)
#pragma unroll
                    for (int n = 0; n < NW; n++) {
                        rp[s][n] = rp2[s][n];
                        rp2[s][n] = *reinterpret_cast<const u32x4 *>(
                            packed + (base + s * 16 + r) * ldp + n2 + (CB / 4) * g + 16 * n);
                    }
            } else {
                DPCA_ZA8_LOAD(nx);
            }
        }
#pragma unroll
        for (int t = 0; t < 4 * NW; t++) {
            // dword q of the A operand: bit pair q of each of the four
            // packed bytes, i.e. samples 4b + q for b = 0..3 (Q8 bytes follow
            // the same order), one shift and mask per four codes
            u32x4 ah[ST], am[ST];
#pragma unroll
            for (int s = 0; s < ST; s++) {
                const uint32_t w = pk[s][t / 4][t % 4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const uint32_t cw = (w >> (2 * q)) & 0x03030303u;
                    ah[s][q] = __builtin_amdgcn_perm(0u, tab[s], cw);
                    am[s][q] = __builtin_amdgcn_perm(0u, MT, cw);
                }
            }
#pragma unroll
            for (int v = 0; v < VT; v++) {
                const u32x4 b = qs[cur][(((CB / 16) * g + t) * VT + v) * 16 + r];
#pragma unroll
                for (int s = 0; s < ST; s++) {
                    acch[s][v] = mfma_i8(ah[s], b, acch[s][v]);
                    if (!slice2_only(v)) accm[s][v] = mfma_i8(am[s], b, accm[s][v]);
                }
            }
        }
        DPCA_ZA8_STORE(cur ^ 1);  // the other buffer: last read before the previous barrier
        __syncthreads();
    }
#undef DPCA_ZA8_STORE
#undef DPCA_ZA8_LOAD
    float *out = Tpart + (int64_t)blockIdx.y * Npad * KP;
#pragma unroll
    for (int oc = 0; oc < CT; oc++) {
        const int col = oc * 16 + r;
        const double sc = slice_scale(__uint_as_float(qmax[blockIdx.y * KP + col]));
#pragma unroll
        for (int s = 0; s < ST; s++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int64_t row = base + s * 16 + 4 * g + q;
                const double a = combine<F, RT>(acch[s], oc, q, r, R, g);
                const double m = combine2<F, RT>(accm[s], oc, q, r, R, g);
                const double mr = mu[row];
                const double delta = mr - rint(63.0 * mr) / 63.0;
                out[row * KP + col] = (float)(dsc[row] * sc * (a / 63.0 - delta * m));
            }
    }
}

// ------------------------------------------------------------- launchers --
// Instantiated tile shapes: (F, RT) of col_tiles(k) for k = 1..64, with the
// sample / SNP tiles per wave chosen so the accumulators fit the registers.
#define DPCA_I8_SHAPES(X) \
    X(1, 0) X(2, 0) X(3, 0) X(4, 0) X(0, 1) X(1, 1) X(2, 1) X(3, 1) X(0, 2) X(1, 2) X(2, 2) X(3, 2)

template <int SA, int F, int RT>
static hipError_t zt_go(hipStream_t st, dim3 grid, const uint8_t *packed, int64_t ldp,
                        const u32x4 *UW8, const unsigned *smax, int64_t rows, int64_t rps,
                        int64_t Mpad, int R, float *part)
{
    if constexpr (SA * (3 * F + RT) > 36) {
        return hipErrorInvalidValue;  // would not fit the registers
    } else {
        k_zt_i8<SA, F, RT><<<grid, 512, 0, st>>>(packed, ldp, UW8, smax, rows, rps, Mpad, R, part);
        return hipGetLastError();
    }
}

template <int ST, int F, int RT, bool SKR>
static hipError_t za_go(hipStream_t st, dim3 grid, const uint8_t *packed, int64_t ldp,
                        const u32x4 *Q8, const unsigned *qmax, const double *mu,
                        const double *dsc, int64_t seg_bytes, int64_t rows, int R, float *Tpart)
{
    if constexpr (ST * (3 * F + RT) > 18) {
        return hipErrorInvalidValue;  // would not fit the registers
    } else {
        k_za_i8<ST, F, RT, SKR><<<grid, 512, 0, st>>>(packed, ldp, Q8, qmax, mu, dsc, seg_bytes, rows,
                                                 R, Tpart);
        return hipGetLastError();
    }
}

// Workgroups of each instance resident per CU, from the runtime (0 where it
// cannot say, e.g. with no device), cached per column-tile shape.
template <typename K>
static int blocks_per_cu(K kern)
{
    int n = 0;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, 512, 0) == hipSuccess ? n : 0;
}

template <int SA, int F, int RT>
static int zt_occ()
{
    if constexpr (SA * (3 * F + RT) > 36) return 0;
    else return blocks_per_cu(k_zt_i8<SA, F, RT>);
}

template <int ST, int F, int RT>
static int za_occ()
{
    if constexpr (ST * (3 * F + RT) > 18) return 0;
    else return blocks_per_cu(k_za_i8<ST, F, RT, false>);
}

struct I8Occ {
    int zt4, zt2, za2, za1;
};

static I8Occ i8_occ(int k)
{
    static std::mutex mu;
    static int dev = -1;
    static I8Occ cache[kMaxKp + 1];
    static bool have[kMaxKp + 1];
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess) return I8Occ{0, 0, 0, 0};
    std::lock_guard<std::mutex> lk(mu);
    if (d != dev) {
        for (bool &h : have) h = false;
        dev = d;
    }
    if (k < 1 || k > kMaxKp) return I8Occ{0, 0, 0, 0};
    if (!have[k]) {
        const ColTiles ct = col_tiles(k);
        I8Occ o{0, 0, 0, 0};
#define X(F_, RT_)                                                                          \
    if (ct.F == F_ && ct.RT == RT_)                                                          \
        o = I8Occ{zt_occ<4, F_, RT_>(), zt_occ<2, F_, RT_>(), za_occ<2, F_, RT_>(),           \
                  za_occ<1, F_, RT_>()};
        DPCA_I8_SHAPES(X)
#undef X
        // an unknown answer is not cached
        if (o.zt4 <= 0 || o.zt2 <= 0 || o.za2 <= 0 || o.za1 <= 0) return o;
        cache[k] = o;
        have[k] = true;
    }
    return cache[k];
}

// Two tiles per wave reuse each operand word from LDS over twice the rows or
// samples of one tile; one tile per wave (two sample tiles for Z^T T) is
// chosen where it doubles the workgroups resident per CU instead.
int zt_i8_tiles(int k, int64_t ldb)
{
    if (col_tiles(k).vt() > 9 || ldb % 128) return 2;
    const I8Occ o = i8_occ(k);
    return (o.zt4 > 0 && o.zt2 >= 2 * o.zt4) ? 2 : 4;
}

int za_i8_tiles(int k)
{
    if (col_tiles(k).vt() > 9) return 1;
    const I8Occ o = i8_occ(k);
    return (o.za2 > 0 && o.za1 >= 2 * o.za2) ? 1 : 2;
}

int zt_i8_blocks_per_cu(int k, int sa_tiles)
{
    const I8Occ o = i8_occ(k);
    return sa_tiles == 4 ? o.zt4 : sa_tiles == 2 ? o.zt2 : 0;
}

// ldp must be a multiple of the tile's byte width (SA*32), rows and
// rows_per_split multiples of 128, rows_per_split at most 2^22 (int32 sums).
hipError_t launch_zt_i8(hipStream_t st, const uint8_t *packed, int64_t rows, int64_t ldp,
                        const void *UW8, const unsigned *smax, int k, int sa_tiles,
                        int64_t rows_per_split, int nsplit, float *part)
{
    const ColTiles ct = col_tiles(k);
    const int64_t tb = (int64_t)sa_tiles * 32;
    if (ldp % tb || rows % 128 || rows_per_split % 128 || rows_per_split > (1 << 22))
        return hipErrorInvalidValue;
    dim3 grid((unsigned)(ldp / tb), (unsigned)nsplit);
    const int64_t Mpad = 4 * ldp;
    const u32x4 *uw = reinterpret_cast<const u32x4 *>(UW8);
#define X(F_, RT_)                                                                          \
    if (ct.F == F_ && ct.RT == RT_) {                                                        \
        if (sa_tiles == 4)                                                                   \
            return zt_go<4, F_, RT_>(st, grid, packed, ldp, uw, smax, rows, rows_per_split,  \
                                     Mpad, ct.R, part);                                      \
        if (sa_tiles == 2)                                                                   \
            return zt_go<2, F_, RT_>(st, grid, packed, ldp, uw, smax, rows, rows_per_split,  \
                                     Mpad, ct.R, part);                                      \
        return hipErrorInvalidValue;                                                         \
    }
    DPCA_I8_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

// rows a multiple of 8*16*st_tiles, seg_bytes of 64, segments of at most
// 65536 samples (int32 sums).
hipError_t launch_za_i8(hipStream_t st, const uint8_t *packed, int64_t rows, int64_t ldp,
                        const void *Q8, const unsigned *qmax, const double *mu,
                        const double *dsc, int k, int st_tiles, int64_t seg_bytes, int nseg,
                        float *Tpart)
{
    const ColTiles ct = col_tiles(k);
    const int64_t per = 8 * 16 * (int64_t)st_tiles;
    // |A| <= 126 per sample: a segment of <= 2^16 samples stays exact in int32
    if (rows % per || seg_bytes % 128 || ldp % 128 || seg_bytes > 16384)
        return hipErrorInvalidValue;
    dim3 grid((unsigned)(rows / per), (unsigned)nseg);
    const u32x4 *q = reinterpret_cast<const u32x4 *>(Q8);
    // the last remainder tile holds only slice 2 when the slices 0 and 1 of
    // the R remainder columns fit the tiles before it
    const bool skr = ct.RT > 0 && 2 * ct.R <= 16 * (ct.RT - 1);
#define X(F_, RT_)                                                                          \
    if (ct.F == F_ && ct.RT == RT_) {                                                        \
        if (st_tiles == 2 && skr)                                                            \
            return za_go<2, F_, RT_, true>(st, grid, packed, ldp, q, qmax, mu, dsc,          \
                                           seg_bytes, rows, ct.R, Tpart);                    \
        if (st_tiles == 2)                                                                   \
            return za_go<2, F_, RT_, false>(st, grid, packed, ldp, q, qmax, mu, dsc,         \
                                            seg_bytes, rows, ct.R, Tpart);                   \
        if (st_tiles == 1 && skr)                                                            \
            return za_go<1, F_, RT_, true>(st, grid, packed, ldp, q, qmax, mu, dsc,          \
                                           seg_bytes, rows, ct.R, Tpart);                    \
        if (st_tiles == 1)                                                                   \
            return za_go<1, F_, RT_, false>(st, grid, packed, ldp, q, qmax, mu, dsc,         \
                                            seg_bytes, rows, ct.R, Tpart);                   \
        return hipErrorInvalidValue;                                                         \
    }
    DPCA_I8_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

}  // namespace dpca
