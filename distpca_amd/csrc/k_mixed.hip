// Mixed-precision form of one pass  A A^T Q = Z^T (Z Q)  on int8 MFMA
// (v_mfma_i32_16x16x64_i8).  The genotype codes enter the products as exact
// int8 values -- h, the allele count (0 when missing), and m, the
// non-missing indicator -- with Z_si = d_s (h_si - mu_s m_si); the FP64
// right-hand operand is split into NS int8 fixed-point slices (NS = 2 or 3)
// with one power-of-two scale per column and group of rows, so every
// product sums exactly in int32 and only the final combination rounds.
//
//   T = Z Q      two int8 rows per SNP (centred h and m) against the slices
//                of Q (scale per sample segment); T_s in FP64 per segment,
//                FP32 partials, segments summed in FP64.
//   Z^T T        two int8 K slots per SNP (h and m) against the slices of
//                U = d T and W = -d mu T (scale per SNP split); splits
//                summed in FP64.
#include "dpca_internal.h"
#include "k_common.h"

#include <algorithm>
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
// An FP64 operand x (Q, or U = d T and W = -d mu T) enters the MFMA as NS
// int8 slices of a fixed-point form with one power-of-two scale s per column
// and group of rows (one workgroup's accumulation range):
// x = s sum_{j<NS} q_j 2^-7j + e, |e| <= s 2^-(7 NS - 6), s >= max|x| / 127
// (NS = 3: |e| <= s 2^-15; NS = 2: s 2^-8).  The slice products sum exactly
// in int32; the slices are combined in FP64 when the group ends.
//
// Column tiles.  The NS slices of the k right-hand columns are laid out as
// 16-wide MFMA B tiles ("virtual tiles", VT = NS F + RT of them): columns
// 16c..16c+15, c < F, take NS tiles each (tile NS c + j holds slice j in
// lane c%16); the R = k - 16F columns left over share RT = ceil(NS R / 16)
// tiles: virtual column w = jR + x holds slice j of column 16F + x, in tile
// NS F + w/16, lane w%16.  So k = 40 takes 8 tiles with 3 slices (not 9)
// and exactly 5 with 2.  Remainders that would need NS tiles anyway are
// padded to whole tiles.
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

// Slice operands reach LDS without passing through registers: one
// global_load_lds_dwordx4 copies 16 bytes per lane from base + off into LDS
// at lds + 16 lane, lds being the wave's (uniform) LDS byte address, set in
// M0.
// The compiler does not see this write, so (a) nothing in the kernel stores
// to the destination buffer through ordinary code, (b) the buffer is read
// only after the issuing waves have waited for their copies (lds_dma_wait)
// and passed a barrier, and (c) it is issued only into a buffer whose
// previous readers have all passed a barrier since.  The instruction has no
// register result; its address operand is read at issue.  M0 is reserved by
// the compiler and not otherwise used by the kernels that call this (no
// message, GWS, indexed-move or compiler-issued LDS-DMA instructions), so it
// is written here without being declared.
// The source is a wave-uniform base (an SGPR pair) plus a 32-bit per-lane
// byte offset, so one offset register serves every copy of a kernel.
__device__ __forceinline__ void lds_dma16(uint32_t lds, const void *base, uint32_t off)
{
    asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, %2" : : "s"(lds), "v"(off), "s"(base) : "memory");
}
// all of this wave's global loads, the LDS copies among them, have landed
__device__ __forceinline__ void lds_dma_wait() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
// all but this wave's N most recently issued global loads have landed (the
// count drains in issue order)
template <int N>
__device__ __forceinline__ void lds_dma_wait_all_but()
{
    static_assert(N >= 0 && N < 64, "vmcnt field");
    asm volatile("s_waitcnt vmcnt(%0)" : : "n"(N) : "memory");
}
// LDS byte address of a __shared__ object (the low half of its generic address)
__device__ __forceinline__ uint32_t lds_addr(const void *p) { return (uint32_t)(uintptr_t)p; }

// power-of-two scale s with max / s < 127 (s = 1 for an all-zero column)
__device__ __forceinline__ double slice_scale(float mx)
{
    int e = 0;
    (void)frexp((double)mx / 127.0, &e);
    return ldexp(1.0, e);
}

// |x| <= 127 s: NS int8 slices of x / s, each residual scaled by 2^7.
template <int NS>
__device__ __forceinline__ void slices(double x, double inv_s, int (&q)[NS])
{
    double v = x * inv_s;              // exact: inv_s is a power of two
#pragma unroll
    for (int j = 0; j < NS; j++) {
        const double a = rint(v);
        q[j] = (int)a;
        v = (v - a) * 128.0;           // exact, |v| <= 64
    }
}

// Slice j and column of virtual column (tile v, lane r); column -1 = empty.
template <int NS>
__device__ __forceinline__ void vcol(int v, int r, int F, int R, int &j, int &col)
{
    if (v < NS * F) {
        j = v % NS;
        col = (v / NS) * 16 + r;
        return;
    }
    const int w = (v - NS * F) * 16 + r;
    j = R ? w / R : NS;
    col = (j < NS) ? 16 * F + w % R : -1;
}

// Value of slice j of column 16F + x for this lane's D rows: the remainder
// tiles' accumulators are spread over lanes 16g + (jR + x) % 16.
template <int NS, int F, int RT>
__device__ __forceinline__ int rem_slice(const i32x4 (&acc)[NS * F + RT], int q, int j, int x,
                                         int R, int g)
{
    const int w = j * R + x;
    const int src = 16 * g + (w & 15);
    int val = 0;
#pragma unroll
    for (int u = 0; u < RT; u++) {
        const int y = __shfl(acc[NS * F + u][q], src, 64);
        if ((w >> 4) == u) val = y;
    }
    return val;
}

// Combined FP64 value sum_{j<NJ} 2^-7j slice_j of output column 16 oc + r for
// D row 4g + q (zero for padding columns).  NJ = NS combines every slice;
// NJ = NS - 1 leaves out the last (a product that needs 2^-7 less precision
// does not compute the tiles that hold only the last slice).
template <int NS, int NJ, int F, int RT>
__device__ __forceinline__ double combine(const i32x4 (&acc)[NS * F + RT], int oc, int q, int r,
                                          int R, int g)
{
    double v = 0.0;
    if (oc < F) {
#pragma unroll
        for (int j = 0; j < NJ; j++) v += ldexp((double)acc[NS * oc + j][q], -7 * j);
        return v;
    }
    // remainder columns: every lane joins the shuffles, lanes r >= R drop out
    const int x = r < R ? r : 0;
#pragma unroll
    for (int j = 0; j < NJ; j++) v += ldexp((double)rem_slice<NS, F, RT>(acc, q, j, x, R, g), -7 * j);
    return r >= R ? 0.0 : v;
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

// One thread per (8-SNP group, column of the virtual tiles): the column's
// NS slices of U and W for the 8 SNPs go to the NS words of their virtual
// tiles; the padding slots of the shared remainder tiles get zero words.
// Columns c < 16 F sit in whole tiles (slice j at tile NS (c / 16) + j, lane
// c % 16); the R columns past them share the remainder tiles (slice j of
// column 16 F + x at virtual column 16 NS F + j R + x).
template <int NS>
__global__ void k_uw8(const double *__restrict__ T, int64_t rows, int kp, int F, int R, int VT,
                      const double *__restrict__ mu, const double *__restrict__ dsc, int64_t rps,
                      const unsigned *__restrict__ smax, u32x4 *__restrict__ UW8)
{
    const int CW = 16 * F + R;                 // columns with slices
    const int PAD = (VT - NS * F) * 16 - NS * R;  // empty slots of the remainder tiles
    const int per = CW + PAD;
    DPCA_EW_LOOP(idx, rows / 8 * per) {
        const int64_t G = idx / per;
        const int c = (int)(idx % per);
        u32x4 *out = UW8 + G * VT * 16;
        if (c >= CW) {  // a padding slot
            const int w = NS * R + (c - CW);
            out[(NS * F + w / 16) * 16 + w % 16] = u32x4{0, 0, 0, 0};
            continue;
        }
        const int64_t split = (G * 8) / rps;
        const double inv = 1.0 / slice_scale(__uint_as_float(smax[split * kp + c]));
        uint32_t wu[NS][2], ww[NS][2];
#pragma unroll
        for (int j = 0; j < NS; j++) wu[j][0] = wu[j][1] = ww[j][0] = ww[j][1] = 0;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int64_t row = G * 8 + e;
            const double t = T[row * kp + c];
            int qu[NS], qw[NS];
            slices<NS>(dsc[row] * t, inv, qu);
            slices<NS>(-(dsc[row] * mu[row]) * t, inv, qw);
            const int sh = 8 * (e & 3);
#pragma unroll
            for (int j = 0; j < NS; j++) {
                wu[j][e >> 2] |= (uint32_t)(qu[j] & 0xFF) << sh;
                ww[j][e >> 2] |= (uint32_t)(qw[j] & 0xFF) << sh;
            }
        }
#pragma unroll
        for (int j = 0; j < NS; j++) {
            int v, r;
            if (c < 16 * F) {
                v = (c / 16) * NS + j;
                r = c % 16;
            } else {
                const int w = j * R + (c - 16 * F);
                v = NS * F + w / 16;
                r = w % 16;
            }
            out[v * 16 + r] = u32x4{wu[j][0], wu[j][1], ww[j][0], ww[j][1]};
        }
    }
}

hipError_t launch_uw8(hipStream_t st, const double *T, int64_t rows, int k, int ns, int kp,
                      const double *mu, const double *dsc, int64_t rps, unsigned *smax, void *UW8,
                      bool have_max)
{
    if (ns != 2 && ns != 3) return hipErrorInvalidValue;
    const ColTiles ct = col_tiles(k, ns);
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
    const int64_t n = rows / 8 * ((ct.vt() - ns * ct.F) * 16 + 16 * ct.F + ct.R - ns * ct.R);
    if (ns == 2)
        k_uw8<2><<<ew_blocks(n), 256, 0, st>>>(T, rows, kp, ct.F, ct.R, ct.vt(), mu, dsc, rps, smax,
                                                reinterpret_cast<u32x4 *>(UW8));
    else
        k_uw8<3><<<ew_blocks(n), 256, 0, st>>>(T, rows, kp, ct.F, ct.R, ct.vt(), mu, dsc, rps, smax,
                                                reinterpret_cast<u32x4 *>(UW8));
    return hipGetLastError();
}

// Tile-order image of the packed matrix for k_zt_i8 (IMG form): for each
// 128-SNP chunk c and TB-byte sample tile x, one block of TB rows (byte
// columns) x 32 dwords, block index c * (ldp / TB) + x; dword `col` of row b
// holds byte column x TB + b of SNPs 4G..4G+3 (byte j = SNP 4G + j) with
// G = col ^ 8 (b & 3).  That is the kernel's LDS tile itself, so a chunk
// reaches LDS as one contiguous DMA.  The swizzle spreads the 4 rows x 4
// pairs of dwords one paired read touches over distinct banks.
// One workgroup (256 threads) per block: the block's 128 x TB input bytes
// are staged in LDS by coalesced row loads, then written in block order.
template <int TB>
__global__ __launch_bounds__(256) void k_zt_image(const uint8_t *__restrict__ packed, int64_t ldp,
                                                  int64_t nblk, uint32_t *__restrict__ img)
{
    __shared__ uint32_t in[128][TB / 4 + 1];
    const int tid = threadIdx.x;
    const int64_t tiles = ldp / TB;
    for (int64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
        const int64_t c = blk / tiles, x = blk % tiles;
        const uint8_t *src = packed + c * 128 * ldp + x * TB;
        __syncthreads();  // the previous block's reads of `in` are done
        for (int i = tid; i < 128 * (TB / 4); i += 256) {
            const int row = i / (TB / 4), q = i % (TB / 4);
            in[row][q] = *reinterpret_cast<const uint32_t *>(src + row * ldp + 4 * q);
        }
        __syncthreads();
        uint32_t *dst = img + blk * (TB * 32);
        for (int o = tid; o < TB * 32; o += 256) {
            const int b = o / 32, col = o % 32;
            const int G = col ^ (8 * (b & 3));
            uint32_t w = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) w |= ((in[4 * G + j][b / 4] >> (8 * (b & 3))) & 0xFFu) << (8 * j);
            dst[o] = w;
        }
    }
}

// Workgroup: WAVES waves x SA tiles x 16 samples (8 x 4: 512 samples = 128
// packed bytes per row) and one SNP split (blockIdx.y); 128-SNP chunks of
// packed bytes staged in LDS transposed to [byte column][4-SNP group], so a
// lane reads the codes of 4 consecutive SNPs for its sample with one dword
// load, and the chunk's UW8 words copied to LDS.  IMG: the chunk's codes come
// from the tile-order image (img, above) by DMA, with no register staging.
template <int NS, int SA, int F, int RT, int WAVES, bool IMG>
__global__ __launch_bounds__(64 * WAVES) void k_zt_i8(const uint8_t *__restrict__ packed,
                                               const uint8_t *__restrict__ img, int64_t ldp,
                                               const u32x4 *__restrict__ UW8,
                                               const unsigned *__restrict__ smax, int64_t rows,
                                               int64_t rows_per_split, int64_t Mpad, int R,
                                               float *__restrict__ part)
{
    constexpr int VT = NS * F + RT;          // virtual column tiles
    constexpr int CT = F + (RT > 0 ? 1 : 0); // output column tiles
    constexpr int KP = CT * 16;
    constexpr int NT = 64 * WAVES;         // threads
    constexpr int TB = SA * 16 * WAVES / 4;  // packed bytes per row in this tile
    constexpr int SC = 128;                // SNPs per chunk
    constexpr int NUWT = (SC / 8) * VT * 16;  // UW8 words per chunk
    // two buffers (three in most IMG forms, below): chunk i is multiplied
    // out of one while chunk i+1 lands in the other (one barrier per chunk).
    // pk rows are byte columns, their dwords 4-SNP groups, stored at column
    // sg ^ (row / 4) in rows of 36 dwords: both the transposing stores (one
    // group, many rows per half-wave) and the paired reads (4 rows x 2
    // groups) then touch distinct banks.  Groups 2x, 2x+1 stay one aligned
    // 8-byte pair, in swapped order when row / 4 is odd.  IMG: rows of 32
    // dwords in the image's order (pairs never swapped).
    constexpr int PKW = IMG ? SC / 4 : SC / 4 + 4;
    // IMG with 4 sample tiles per wave and 4 or more column tiles (the
    // registers hold one workgroup per CU anyway), where three buffers fit
    // the LDS: three buffers, chunk i+2's copies in flight while chunk i is
    // multiplied (twice the time to cover HBM latency; 0.5-1.2% faster)
    constexpr int BUFB = TB * PKW * 4 + NUWT * 16;  // LDS bytes per buffer
    constexpr int NBUF = (IMG && SA >= 4 && VT >= 4 && 3 * BUFB <= 160 * 1024) ? 3 : 2;
    static_assert(SA % 2 == 0, "row / 4 parity must follow the tile index");
    __shared__ uint32_t pk[NBUF][TB][PKW];
    __shared__ u32x4 uw[NBUF][NUWT];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int r = lane & 15, g = lane >> 4;
    int bx, by;  // sample tile, SNP split
    xcd_tile((int)(ldp / TB), bx, by);
    const int64_t mb = (int64_t)bx * TB;
    const int64_t sbeg = (int64_t)by * rows_per_split;
    int64_t send = sbeg + rows_per_split;
    if (send > rows) send = rows;
    const uint32_t HT = 0x02010000u, MT = 0x01010001u;  // h, m by 2-bit code

    i32x4 acc[SA][VT];
#pragma unroll
    for (int a = 0; a < SA; a++)
#pragma unroll
        for (int v = 0; v < VT; v++) acc[a][v] = i32x4{0, 0, 0, 0};

    constexpr int NPK = (SC / 4) * (TB / 4) / NT;   // 4x4 byte blocks per thread
    constexpr int NUW = (NUWT + NT - 1) / NT;       // UW8 words per thread
    static_assert((SC / 4) * (TB / 4) % NT == 0, "tile");
    // software pipeline: while the current chunk is multiplied, the next
    // chunk's packed codes are loaded into registers (stored into the other
    // LDS buffer, transposed, at the end of the chunk) and its UW8 words are
    // copied straight into the other buffer (lds_dma16)
    uint32_t rpk[NPK][4];
    // each thread's byte offset within a chunk (32 bits: < 32 rows of ldp
    // bytes); the chunk's own offset is uniform, so the address arithmetic of
    // a chunk's loads is scalar work plus one add
    uint32_t poff[NPK];
#pragma unroll
    for (int u = 0; u < NPK; u++) {
        const int it = tid + NT * u;
        poff[u] = (uint32_t)((4 * (it / (TB / 4))) * ldp + mb + 4 * (it % (TB / 4)));
    }
#define DPCA_ZT8_LOADP(c0, dst)                                                             \
    do {                                                                                    \
        _Pragma("unroll") for (int j = 0; j < 4; j++) {                                    \
            const uint8_t *b_ = packed + ((c0) + j) * ldp;                                  \
            _Pragma("unroll") for (int u = 0; u < NPK; u++) dst[u][j] =                    \
                *reinterpret_cast<const uint32_t *>(b_ + poff[u]);                          \
        }                                                                                   \
    } while (0)
    // lane l of wave w copies word NT u + 64 w + l of the chunk's UW8 words
    // (NUWT is a multiple of 64, so each wave's test is uniform).  Three
    // buffers: every wave issues NUW copies (one past the end copies word
    // w mod NUWT again, the same bytes to the same place), so that its count
    // of copies in flight is a constant.
    const uint32_t uw_lds = lds_addr(&uw[0][0]);
    const int wave_u = __builtin_amdgcn_readfirstlane(wave);
    __builtin_assume(wave_u >= 0 && wave_u < WAVES);
    const uint32_t loff = 16u * (uint32_t)lane;  // per-lane source offset of a copy
#define DPCA_ZT8_LOADU(c0, buf)                                                             \
    do {                                                                                    \
        const u32x4 *q_ = UW8 + ((c0) / 8) * VT * 16;                                       \
        _Pragma("unroll") for (int u = 0; u < NUW; u++) {                                  \
            int w_ = NT * u + 64 * wave_u;                                                  \
            if (NBUF == 3) w_ %= NUWT;                                                      \
            if (w_ < NUWT)                                                                  \
                lds_dma16(uw_lds + 16u * (uint32_t)((buf) * NUWT + w_), q_ + w_, loff);     \
        }                                                                                   \
    } while (0)
    // 4x4 byte transpose of each block into [byte column][SNP group]
#define DPCA_ZT8_STORE(buf)                                                                 \
    do {                                                                                    \
        _Pragma("unroll") for (int u = 0; u < NPK; u++) {                                  \
            const int it_ = tid + NT * u;                                                   \
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
    } while (0)
    static_assert(NUWT % 64 == 0, "whole waves of UW8 words");
    // IMG: a chunk's block of the image (TB rows x 32 dwords) into pk[buf]
    constexpr int NIW = TB * SC / 16;  // 16-byte words per block (TB x SC bytes)
    static_assert(NIW % 64 == 0, "whole waves of image words");
    const uint32_t pk_lds = lds_addr(&pk[0][0][0]);
#define DPCA_ZT8_LOADI(c0, buf)                                                             \
    do {                                                                                    \
        const uint8_t *b_ = img + (((c0) / SC) * (ldp / TB) + bx) * (int64_t)(TB * SC);     \
        _Pragma("unroll") for (int u = 0; u < (NIW + NT - 1) / NT; u++) {                  \
            const int w_ = NT * u + 64 * wave_u;                                            \
            if (w_ < NIW)                                                                   \
                lds_dma16(pk_lds + 16u * (uint32_t)((buf) * NIW + w_), b_ + 16 * w_, loff);    \
        }                                                                                   \
    } while (0)
    // copies per wave and chunk (three buffers: both counts uniform)
    constexpr int NCP = (NIW + NT - 1) / NT + NUW;
    static_assert(NBUF == 2 || NIW % NT == 0, "uniform image copies");
    {
        const int64_t c0 = sbeg < send ? sbeg : 0;
        if constexpr (IMG) {
            DPCA_ZT8_LOADI(c0, 0);
        } else {
            DPCA_ZT8_LOADP(c0, rpk);
        }
        DPCA_ZT8_LOADU(c0, 0);
        if constexpr (!IMG) DPCA_ZT8_STORE(0);
        if constexpr (NBUF == 3) {
            const int64_t c1 = (c0 + SC < send) ? c0 + SC : c0;
            DPCA_ZT8_LOADI(c1, 1);
            DPCA_ZT8_LOADU(c1, 1);
        }
    }
    if constexpr (NBUF == 3) lds_dma_wait_all_but<NCP>();  // chunk 1 may be in flight
    else lds_dma_wait();
    __syncthreads();

    int cur = 0;
    for (int64_t s0 = sbeg; s0 < send; s0 += SC, cur = (cur + 1 == NBUF) ? 0 : cur + 1) {
        if constexpr (NBUF == 3) {
            // chunk i+2 (the last chunk again near the end) into the buffer
            // chunk i-1 was read from before the previous barrier
            const int64_t n2 = (s0 + 2 * SC < send) ? s0 + 2 * SC : send - SC;
            const int nb = (cur == 0) ? 2 : cur - 1;
            DPCA_ZT8_LOADI(n2, nb);
            DPCA_ZT8_LOADU(n2, nb);
        } else {
            // unconditional (re-reads the last chunk at the end)
            const int64_t nx = (s0 + SC < send) ? s0 + SC : s0;
            if constexpr (IMG) {
                DPCA_ZT8_LOADI(nx, cur ^ 1);  // last read before the previous barrier
            } else {
                DPCA_ZT8_LOADP(nx, rpk);
            }
            DPCA_ZT8_LOADU(nx, cur ^ 1);
        }
        // The chunk's B tiles (K step t, tile v) form one sequence i = t VT + v
        // read from LDS ZB tiles ahead of their MFMAs, and the packed words
        // of K step t + 1 are read while step t multiplies, so the LDS
        // latency hides behind the matrix cores instead of stalling them.
        constexpr int NKS = SC / 32;            // K steps per chunk
        constexpr int NB = NKS * VT;            // B tiles per chunk
        // B tiles in flight: 12 with 8-wave workgroups (measured at 100k x
        // 1M, nrhs 40: 4 -> 8 -> 12 -> 16 tiles, 16.0 -> 15.6 -> 15.1 ->
        // 15.2 ms; no spills), 4 in the 4-wave tuning forms
        constexpr int ZB = WAVES == 8 ? (NB < 12 ? NB : 12) : 4;
        const u32x4 *ub = &uw[cur][g * VT * 16 + r];
        auto boff = [](int i) constexpr { return ((i / VT) * 4 * VT + i % VT) * 16; };
        u32x4 bq[ZB];
#pragma unroll
        for (int i = 0; i < ZB; i++) bq[i] = ub[boff(i)];
        uint2 wq[SA];
        auto read_words = [&](int t) {
#pragma unroll
            for (int a = 0; a < SA; a++) {
                const int i = wave * SA * 16 + a * 16 + r;  // sample within the tile
                if constexpr (IMG) {
                    wq[a] = *reinterpret_cast<const uint2 *>(
                        &pk[cur][i >> 2][(8 * t + 2 * g) ^ (8 * ((i >> 2) & 3))]);
                } else {
                    const int sw = (wave * SA + a) & 30;     // (i / 16) & ~1
                    wq[a] = *reinterpret_cast<const uint2 *>(&pk[cur][i >> 2][(8 * t + 2 * g) ^ sw]);
                }
            }
        };
        // A operands of K step t + 1 are built from its packed words (read at
        // the start of step t) while step t multiplies, one sample tile per B
        // tile from the third on (the last tile when there are fewer), so the
        // VALU work interleaves with the MFMAs
        u32x4 av[2][SA];
        auto build = [&](int t, int a) {
            const int sh = 2 * (r & 3);                      // sample i & 3 = r & 3
            const bool sw = !IMG && (a & 1);
            const uint32_t x0 = sw ? wq[a].y : wq[a].x;   // group 8t + 2g
            const uint32_t x1 = sw ? wq[a].x : wq[a].y;   // group 8t + 2g + 1
            const uint32_t c0 = (x0 >> sh) & 0x03030303u;
            const uint32_t c1 = (x1 >> sh) & 0x03030303u;
            av[t & 1][a] = u32x4{__builtin_amdgcn_perm(0u, HT, c0), __builtin_amdgcn_perm(0u, HT, c1),
                                 __builtin_amdgcn_perm(0u, MT, c0), __builtin_amdgcn_perm(0u, MT, c1)};
        };
        read_words(0);
#pragma unroll
        for (int a = 0; a < SA; a++) build(0, a);
#pragma unroll
        for (int i = 0; i < NB; i++) {
            const int t = i / VT, v = i % VT;
            if (v == 0 && t + 1 < NKS) read_words(t + 1);
            const u32x4 b = bq[i % ZB];
            if (i + ZB < NB) bq[i % ZB] = ub[boff(i + ZB)];
#pragma unroll
            for (int a = 0; a < SA; a++) acc[a][v] = mfma_i8(av[t & 1][a], b, acc[a][v]);
            if (t + 1 < NKS) {
#pragma unroll
                for (int a = 0; a < SA; a++)
                    if ((VT >= 3 ? 2 + a * (VT - 2) / SA : VT - 1) == v) build(t + 1, a);
            }
            // keep each tile's LDS reads, MFMAs and A piece together (the
            // scheduler would otherwise sink the reads next to their use and
            // bunch the VALU work); only SALU moves across
            __builtin_amdgcn_sched_barrier(0x4);
        }
        if constexpr (!IMG) DPCA_ZT8_STORE(cur ^ 1);  // last read before the previous barrier
        if constexpr (NBUF == 3) lds_dma_wait_all_but<NCP>();  // chunk i+1 landed
        else lds_dma_wait();
        __syncthreads();
    }
    lds_dma_wait();  // no copy into this workgroup's LDS outlives it
#undef DPCA_ZT8_STORE
#undef DPCA_ZT8_LOADI
#undef DPCA_ZT8_LOADU
#undef DPCA_ZT8_LOADP
    float *out = part + ((int64_t)by * Mpad + 4 * mb + wave * SA * 16) * KP;
#pragma unroll
    for (int oc = 0; oc < CT; oc++) {
        const int col = oc * 16 + r;
        const double s = slice_scale(__uint_as_float(smax[by * KP + col]));
#pragma unroll
        for (int a = 0; a < SA; a++)
#pragma unroll
            for (int q = 0; q < 4; q++)
                out[(a * 16 + 4 * g + q) * KP + col] =
                    (float)(s * combine<NS, NS, F, RT>(acc[a], oc, q, r, R, g));
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
template <int NS>
__global__ void k_q8(const double *__restrict__ Q, int64_t Mpad, int kp, int F, int R, int VT,
                     int64_t seg_samples, const unsigned *__restrict__ qmax,
                     u32x4 *__restrict__ Q8)
{
    DPCA_EW_LOOP(idx, Mpad / 16 * VT * 16) {
        const int64_t G = idx / (VT * 16);
        const int v = (int)(idx % (VT * 16)) / 16, r = (int)(idx % 16);
        int j, c;
        vcol<NS>(v, r, F, R, j, c);
        uint32_t w[4] = {0, 0, 0, 0};
        if (c >= 0) {
            const int64_t seg = (G * 16) / seg_samples;
            const double inv = 1.0 / slice_scale(__uint_as_float(qmax[seg * kp + c]));
#pragma unroll
            for (int e = 0; e < 16; e++) {
                int q[NS];
                slices<NS>(Q[(G * 16 + 4 * (e & 3) + (e >> 2)) * kp + c], inv, q);
                w[e >> 2] |= (uint32_t)(q[j] & 0xFF) << (8 * (e & 3));
            }
        }
        Q8[idx] = u32x4{w[0], w[1], w[2], w[3]};
    }
}

hipError_t launch_q8(hipStream_t st, const double *Q, int64_t Mpad, int k, int ns, int kp,
                     int64_t seg_bytes, unsigned *qmax, void *Q8)
{
    if (ns != 2 && ns != 3) return hipErrorInvalidValue;
    const ColTiles ct = col_tiles(k, ns);
    const int64_t segs = 4 * seg_bytes;  // samples per segment
    if (Mpad % 256 || segs % 256 || kp * 4 > 256 || ct.kp() != kp) return hipErrorInvalidValue;
    const int64_t nseg = (Mpad + segs - 1) / segs;
    hipError_t e = launch_fill_u8(st, reinterpret_cast<uint8_t *>(qmax), nseg * kp * 4, 0);
    if (e != hipSuccess) return e;
    dim3 g1((unsigned)((segs + 127) / 128), (unsigned)nseg);
    k_split_max<<<g1, 4 * kp, 0, st>>>(Q, Mpad, kp, nullptr, nullptr, segs, qmax);
    const int64_t n = Mpad / 16 * ct.vt() * 16;
    if (ns == 2)
        k_q8<2><<<ew_blocks(n), 256, 0, st>>>(Q, Mpad, kp, ct.F, ct.R, ct.vt(), segs, qmax,
                                               reinterpret_cast<u32x4 *>(Q8));
    else
        k_q8<3><<<ew_blocks(n), 256, 0, st>>>(Q, Mpad, kp, ct.F, ct.R, ct.vt(), segs, qmax,
                                               reinterpret_cast<u32x4 *>(Q8));
    return hipGetLastError();
}

// Workgroup: WAVES waves x ST tiles x 16 SNPs, one sample segment (blockIdx.y).
// Per chunk of CB packed bytes per row (4 CB samples), lane group g owns the
// CB samples of its CB / 4 packed bytes; at K step t it supplies samples
// CB g + 16t .. +15.  The chunk's Q8 words (CDMA, below: and its codes) are
// copied into LDS.
template <int NS, int ST, int F, int RT, bool SKR, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_za_i8(const uint8_t *__restrict__ packed, int64_t ldp,
                                               const u32x4 *__restrict__ Q8,
                                               const unsigned *__restrict__ qmax,
                                               const double *__restrict__ mu,
                                               const double *__restrict__ dsc, int64_t seg_bytes,
                                               int64_t Npad, int R, float *__restrict__ Tpart)
{
    constexpr int VT = NS * F + RT;
    constexpr int CT = F + (RT > 0 ? 1 : 0);
    constexpr int KP = CT * 16;
    // CDMA (8 waves, 5..8 column tiles: the resident workgroups per CU are
    // set by the registers anyway): chunks of 64 packed bytes, the codes
    // copied into LDS beside the Q8 words, three buffers with chunk i+2's
    // copies in flight while chunk i is multiplied (100k x 1M: 3 slices, 8
    // tiles, 14.9 -> 14.4 ms; 2 slices, 5 tiles, 10.0 -> 9.9 ms; at 4 tiles
    // no change).  Otherwise packed bytes per chunk: 128 where two LDS
    // buffers of its Q8 words (CB * VT * 64 bytes each) fit the 160 KB of
    // LDS, else 64, and the codes loaded into registers a chunk ahead.
    constexpr bool CDMA = WAVES == 8 && VT >= 5 && VT <= 8;
    constexpr int CB = CDMA ? 64 : (VT <= 10) ? 128 : 64;
    constexpr int NBUF = CDMA ? 3 : 2;
    constexpr int NW = CB / 64;                 // 16-byte words per lane and row
    constexpr int NQT = (CB / 4) * VT * 16;     // Q8 words per chunk
    constexpr int NT = 64 * WAVES;              // threads
    constexpr int NQ = (NQT + NT - 1) / NT;
    constexpr int ROWS = WAVES * ST * 16;       // SNP rows per workgroup
    constexpr int NCT = CDMA ? ROWS * (CB / 16) : 1;  // code words per chunk
    __shared__ u32x4 qs[NBUF][NQT];            // one barrier per chunk
    __shared__ u32x4 cs[CDMA ? NBUF : 1][NCT];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int r = lane & 15, g = lane >> 4;
    int bx, by;  // SNP row block, sample segment
    xcd_tile((int)(Npad / (WAVES * ST * 16)), bx, by);
    const int64_t base = (int64_t)bx * (WAVES * ST * 16) + wave * (ST * 16);
    const int64_t b0 = (int64_t)by * seg_bytes;
    int64_t b1 = b0 + seg_bytes;
    if (b1 > ldp) b1 = ldp;
    const uint32_t MT = 0x01010001u;  // m by 2-bit code
    // Per-row A values 63 h - mh with mh = rint(63 mu) (exact int8, 0 when
    // missing), as a byte table by 2-bit code: T_s = d_s (sum (63 h - mh m) Q
    // / 63 - delta_s sum m Q), delta_s = mu_s - mh / 63, |delta_s| <= 1/126,
    // so the mask product needs one slice fewer (2^-7 less precision).
    uint32_t tab[ST];
#pragma unroll
    for (int s = 0; s < ST; s++) {
        const int mh = (int)rint(63.0 * mu[base + s * 16 + r]);  // mu in [0, 2]
        tab[s] = ((uint32_t)(-mh) & 0xFFu) | (((uint32_t)(63 - mh) & 0xFFu) << 16) |
                 (((uint32_t)(126 - mh) & 0xFFu) << 24);
    }
    // tiles holding only the last slice (of whole columns, and the last
    // remainder tile when SKR) are left out of the mask product
    auto last_only = [](int v) constexpr {
        return v < NS * F ? v % NS == NS - 1 : (SKR && v == NS * F + RT - 1);
    };

    i32x4 acch[ST][VT], accm[ST][VT];
#pragma unroll
    for (int s = 0; s < ST; s++)
#pragma unroll
        for (int v = 0; v < VT; v++) {
            acch[s][v] = i32x4{0, 0, 0, 0};
            accm[s][v] = i32x4{0, 0, 0, 0};
        }

    // software pipeline: while the current chunk is multiplied, the next
    // chunk's packed bytes are loaded into registers and its Q8 words copied
    // straight into the other LDS buffer (lds_dma16; lane l of wave w copies
    // word NT u + 64 w + l, NQT being a multiple of 64)
    static_assert(NQT % 64 == 0, "whole waves of Q8 words");
    const uint32_t qs_lds = lds_addr(&qs[0][0]);
    const int wave_u = __builtin_amdgcn_readfirstlane(wave);
    __builtin_assume(wave_u >= 0 && wave_u < WAVES);
    const uint32_t loff = 16u * (uint32_t)lane;  // per-lane source offset of a Q8 copy
    u32x4 rp[ST][NW];
    // CDMA: every wave issues NQ copies (one past the end copies word w mod
    // NQT again, the same bytes to the same place), so that its count of
    // copies in flight is a constant
#define DPCA_ZA8_LOADQ(c0, buf)                                                             \
    do {                                                                                    \
        const u32x4 *src_ = Q8 + ((c0) / 4) * VT * 16;                                      \
        _Pragma("unroll") for (int u = 0; u < NQ; u++) {                                   \
            int w_ = NT * u + 64 * wave_u;                                                  \
            if (CDMA) w_ %= NQT;                                                            \
            if (w_ < NQT)                                                                   \
                lds_dma16(qs_lds + 16u * (uint32_t)((buf) * NQT + w_), src_ + w_, loff);    \
        }                                                                                   \
    } while (0)
    // CDMA: code word w of a chunk is bytes 16 (w % 4) .. +15 of the chunk in
    // the workgroup's row w / 4, at cs[buf][w]
    constexpr int NCW = CDMA ? NCT / NT : 0;    // code copies per wave
    static_assert(!CDMA || NCT % NT == 0, "uniform code copies");
    const uint32_t cs_lds = lds_addr(&cs[0][0]);
    const uint8_t *cblk = packed + (int64_t)bx * ROWS * ldp;
    // per-lane offset of a code copy (64 | w, so row w / 4 + lane / 4, bytes
    // 16 (lane % 4)); < 2^32: the workgroup's rows span ROWS x ldp bytes
    const uint32_t coff = (uint32_t)((lane / 4) * ldp + 16 * (lane % 4));
#define DPCA_ZA8_LOADC(c0, buf)                                                             \
    do {                                                                                    \
        _Pragma("unroll") for (int u = 0; u < NCW; u++) {                                  \
            const int w_ = NT * u + 64 * wave_u;                                            \
            lds_dma16(cs_lds + 16u * (uint32_t)((buf) * NCT + w_),                          \
                      cblk + (int64_t)(w_ / 4) * ldp + (c0), coff);                         \
        }                                                                                   \
    } while (0)
#define DPCA_ZA8_LOADP(c0)                                                                  \
    do {                                                                                    \
        _Pragma("unroll") for (int s = 0; s < ST; s++)                                     \
            _Pragma("unroll") for (int n = 0; n < NW; n++) rp[s][n] =                      \
                *reinterpret_cast<const u32x4 *>(packed + (base + s * 16 + r) * ldp + (c0) + \
                                                 (CB / 4) * g + 16 * n);                    \
    } while (0)
    DPCA_ZA8_LOADQ(b0 < b1 ? b0 : 0, 0);
    if constexpr (CDMA) {
        const int64_t c0 = b0 < b1 ? b0 : 0;
        const int64_t c1 = (c0 + CB < b1) ? c0 + CB : c0;
        DPCA_ZA8_LOADC(c0, 0);
        DPCA_ZA8_LOADQ(c1, 1);
        DPCA_ZA8_LOADC(c1, 1);
    } else {
        DPCA_ZA8_LOADP(b0 < b1 ? b0 : 0);
    }
    // where the registers allow, the packed codes (HBM) run two chunks
    // ahead and the Q8 words (L2) one
    constexpr bool PF2 = !CDMA && ST * VT < 16;
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
    if constexpr (CDMA) lds_dma_wait_all_but<NQ + NCW>();  // chunk 1 may be in flight
    else lds_dma_wait();
    __syncthreads();

    int cur = 0;
    for (int64_t cb = b0; cb < b1; cb += CB, cur = (cur + 1 == NBUF) ? 0 : cur + 1) {
        u32x4 pk[ST][NW];
        if constexpr (CDMA) {
            // chunk i+2 (the last chunk again near the end) into the buffer
            // chunk i-1 was read from before the previous barrier
            const int64_t n2 = (cb + 2 * CB < b1) ? cb + 2 * CB : b1 - CB;
            const int nb = (cur == 0) ? 2 : cur - 1;
            DPCA_ZA8_LOADQ(n2, nb);
            DPCA_ZA8_LOADC(n2, nb);
#pragma unroll
            for (int s = 0; s < ST; s++)
                pk[s][0] = cs[cur][(wave * ST * 16 + s * 16 + r) * (CB / 16) + g];
        } else {
#pragma unroll
            for (int s = 0; s < ST; s++)
#pragma unroll
                for (int n = 0; n < NW; n++) pk[s][n] = rp[s][n];
        }
        if constexpr (!CDMA) {
            // unconditional (re-reads the last chunk at the end)
            const int64_t nx = (cb + CB < b1) ? cb + CB : cb;
            DPCA_ZA8_LOADQ(nx, cur ^ 1);  // last read before the previous barrier
            if constexpr (PF2) {
                const int64_t n2 = (cb + 2 * CB < b1) ? cb + 2 * CB : nx;
#pragma unroll
                for (int s = 0; s < ST; s++)
#pragma unroll
                    for (int n = 0; n < NW; n++) {
                        rp[s][n] = rp2[s][n];
                        rp2[s][n] = *reinterpret_cast<const u32x4 *>(
                            packed + (base + s * 16 + r) * ldp + n2 + (CB / 4) * g + 16 * n);
                    }
            } else {
                DPCA_ZA8_LOADP(nx);
            }
        }
        // The chunk's B tiles (K step t, tile v) form one sequence i = t VT + v
        // read from LDS ZB tiles ahead of their MFMAs (the LDS latency then
        // hides behind the matrix cores instead of stalling them).
        constexpr int NB = 4 * NW * VT;         // B tiles per chunk
        // B tiles in flight: 12 with 8-wave workgroups of 6 or more column
        // tiles (14.5 -> 14.0-14.1 ms at 100k x 1M, nrhs 40; no spills), 4
        // with fewer (20k x 500k, nrhs 20, 4 tiles: 0.97 ms against 1.01 ms
        // with 12) and in the 4-wave tuning forms
        constexpr int ZB = (WAVES == 8 && VT >= 6) ? (NB < 12 ? NB : 12) : 4;
        const u32x4 *qb = &qs[cur][(CB / 16) * g * VT * 16 + r];
        u32x4 bq[ZB];
#pragma unroll
        for (int i = 0; i < ZB; i++) bq[i] = qb[i * 16];
        // A operands of K step t + 1 are built while step t multiplies: its
        // 4 ST (tile, dword) pieces are spread over the step's B tiles, so
        // the VALU work interleaves with the MFMAs of the same wave.
        // dword q of an A operand: bit pair q of each of the four packed
        // bytes, i.e. samples 4b + q for b = 0..3 (Q8 bytes follow the same
        // order), one shift and mask per four codes.
        constexpr int NKS = 4 * NW;             // K steps per chunk
        constexpr int NP = 4 * ST;              // A pieces per K step
        u32x4 ah[2][ST], am[2][ST];
        auto build = [&](int t, int p) {
            const int s = p / 4, q = p % 4;
            const uint32_t cw = (pk[s][t / 4][t % 4] >> (2 * q)) & 0x03030303u;
            ah[t & 1][s][q] = __builtin_amdgcn_perm(0u, tab[s], cw);
            am[t & 1][s][q] = __builtin_amdgcn_perm(0u, MT, cw);
        };
#pragma unroll
        for (int p = 0; p < NP; p++) build(0, p);
#pragma unroll
        for (int i = 0; i < NB; i++) {
            const int t = i / VT, v = i % VT;
            const u32x4 b = bq[i % ZB];
            if (i + ZB < NB) bq[i % ZB] = qb[(i + ZB) * 16];
#pragma unroll
            for (int s = 0; s < ST; s++) {
                acch[s][v] = mfma_i8(ah[t & 1][s], b, acch[s][v]);
                if (!last_only(v)) accm[s][v] = mfma_i8(am[t & 1][s], b, accm[s][v]);
            }
            if (t + 1 < NKS) {
#pragma unroll
                for (int p = 0; p < NP; p++)
                    if (p * VT / NP == v) build(t + 1, p);
            }
            // keep each tile's LDS read, MFMAs and A piece together (the
            // scheduler would otherwise sink the reads next to their use and
            // bunch the VALU work); only SALU moves across
            __builtin_amdgcn_sched_barrier(0x4);
        }
        if constexpr (CDMA) lds_dma_wait_all_but<NQ + NCW>();  // chunk i+1 landed
        else lds_dma_wait();
        __syncthreads();
    }
    lds_dma_wait();  // no copy into this workgroup's LDS outlives it
#undef DPCA_ZA8_LOADC
#undef DPCA_ZA8_LOADP
#undef DPCA_ZA8_LOADQ
    float *out = Tpart + (int64_t)by * Npad * KP;
#pragma unroll
    for (int oc = 0; oc < CT; oc++) {
        const int col = oc * 16 + r;
        const double sc = slice_scale(__uint_as_float(qmax[by * KP + col]));
#pragma unroll
        for (int s = 0; s < ST; s++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int64_t row = base + s * 16 + 4 * g + q;
                const double a = combine<NS, NS, F, RT>(acch[s], oc, q, r, R, g);
                const double m = combine<NS, NS - 1, F, RT>(accm[s], oc, q, r, R, g);
                const double mr = mu[row];
                const double delta = mr - rint(63.0 * mr) / 63.0;
                out[row * KP + col] = (float)(dsc[row] * sc * (a / 63.0 - delta * m));
            }
    }
}

// ------------------------------------------------------------- launchers --
// Instantiated tile shapes: (NS, F, RT) of col_tiles(k, NS) for k = 1..64 and
// NS = 3 or 2, with the sample / SNP tiles per wave whose accumulators,
// double-buffered A operands and prefetch registers fit without spilling
// (zt_fits / za_fits, checked against the compiler's register report).
#define DPCA_I8_SHAPES(X)                                                                  \
    X(3, 1, 0) X(3, 2, 0) X(3, 3, 0) X(3, 4, 0) X(3, 0, 1) X(3, 1, 1) X(3, 2, 1) X(3, 3, 1)   \
        X(3, 0, 2) X(3, 1, 2) X(3, 2, 2) X(3, 3, 2) X(2, 1, 0) X(2, 2, 0) X(2, 3, 0)          \
            X(2, 4, 0) X(2, 0, 1) X(2, 1, 1) X(2, 2, 1) X(2, 3, 1)

// Workgroups of 8 waves (two per SIMD, 256 registers each) or 4 (one per
// SIMD with up to 512 registers, the accumulators in AGPRs: twice the tile
// per wave, so each LDS operand word feeds twice the MFMAs; measured slower
// at 100k x 1M, kept as a tuning form).
constexpr bool zt_fits(int sa, int vt, int waves)
{
    return waves == 8 ? sa <= 4 && sa * vt <= 32 : sa * vt <= 48;
}
constexpr bool za_fits(int st, int vt, int waves)
{
    return waves == 8 ? st <= 2 && st * vt <= 16 : st * vt <= 24;
}
// packed bytes per row of one Z^T T workgroup's sample tile
constexpr int zt_tile_bytes(int sa, int waves) { return 4 * sa * waves; }

template <int NS, int SA, int F, int RT, int W>
static hipError_t zt_go(hipStream_t st, dim3 grid, const uint8_t *packed, const uint8_t *img,
                        int64_t ldp, const u32x4 *UW8, const unsigned *smax, int64_t rows,
                        int64_t rps, int64_t Mpad, int R, float *part)
{
    if constexpr (!zt_fits(SA, NS * F + RT, W)) {
        return hipErrorInvalidValue;  // would not fit the registers
    } else {
        if constexpr (W == 8) {  // the image form exists for 8-wave workgroups
            if (img) {
                k_zt_i8<NS, SA, F, RT, W, true><<<grid, 64 * W, 0, st>>>(
                    packed, img, ldp, UW8, smax, rows, rps, Mpad, R, part);
                return hipGetLastError();
            }
        }
        if (img) return hipErrorInvalidValue;
        k_zt_i8<NS, SA, F, RT, W, false><<<grid, 64 * W, 0, st>>>(packed, nullptr, ldp, UW8, smax,
                                                                 rows, rps, Mpad, R, part);
        return hipGetLastError();
    }
}

template <int NS, int ST, int F, int RT, bool SKR, int W>
static hipError_t za_go(hipStream_t st, dim3 grid, const uint8_t *packed, int64_t ldp,
                        const u32x4 *Q8, const unsigned *qmax, const double *mu,
                        const double *dsc, int64_t seg_bytes, int64_t rows, int R, float *Tpart)
{
    if constexpr (!za_fits(ST, NS * F + RT, W)) {
        return hipErrorInvalidValue;  // would not fit the registers
    } else {
        k_za_i8<NS, ST, F, RT, SKR, W><<<grid, 64 * W, 0, st>>>(packed, ldp, Q8, qmax, mu, dsc,
                                                               seg_bytes, rows, R, Tpart);
        return hipGetLastError();
    }
}

// Workgroups of each instance resident per CU, from the runtime (0 where it
// cannot say, e.g. with no device), cached per shape.
template <typename K>
static int blocks_per_cu(K kern, int threads)
{
    int n = 0;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, threads, 0) == hipSuccess ? n : 0;
}

template <int NS, int SA, int F, int RT, int W>
static int zt_occ()
{
    if constexpr (!zt_fits(SA, NS * F + RT, W)) return 0;
    else return blocks_per_cu(k_zt_i8<NS, SA, F, RT, W, false>, 64 * W);
}

template <int NS, int ST, int F, int RT, int W>
static int za_occ()
{
    if constexpr (!za_fits(ST, NS * F + RT, W)) return 0;
    else return blocks_per_cu(k_za_i8<NS, ST, F, RT, false, W>, 64 * W);
}

// Tile counts per wave, largest first: Z^T T sample tiles, Z Q SNP tiles.
constexpr int kZtTiles[3] = {8, 4, 2};
constexpr int kZaTiles[3] = {4, 2, 1};

struct I8Occ {
    int zt[2][3], za[2][3];  // [waves 8, 4][tiles in kZtTiles / kZaTiles order]
};

static I8Occ i8_occ(int k, int ns)
{
    static std::mutex mu;
    static int dev = -1;
    static I8Occ cache[2][kMaxKp + 1];
    static bool have[2][kMaxKp + 1];
    I8Occ none{};
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess) return none;
    std::lock_guard<std::mutex> lk(mu);
    if (d != dev) {
        for (auto &hv : have)
            for (bool &h : hv) h = false;
        dev = d;
    }
    if (k < 1 || k > kMaxKp || (ns != 2 && ns != 3)) return none;
    const int ni = ns - 2;
    if (!have[ni][k]) {
        const ColTiles ct = col_tiles(k, ns);
        I8Occ o{};
#define OCC_W(NS_, F_, RT_, W_, wi)                                                         \
    o.zt[wi][0] = zt_occ<NS_, 8, F_, RT_, W_>();                                             \
    o.zt[wi][1] = zt_occ<NS_, 4, F_, RT_, W_>();                                             \
    o.zt[wi][2] = zt_occ<NS_, 2, F_, RT_, W_>();                                             \
    o.za[wi][0] = za_occ<NS_, 4, F_, RT_, W_>();                                             \
    o.za[wi][1] = za_occ<NS_, 2, F_, RT_, W_>();                                             \
    o.za[wi][2] = za_occ<NS_, 1, F_, RT_, W_>();
#define X(NS_, F_, RT_)                                                                     \
    if (ct.NS == NS_ && ct.F == F_ && ct.RT == RT_) {                                        \
        OCC_W(NS_, F_, RT_, 8, 0)                                                            \
        OCC_W(NS_, F_, RT_, 4, 1)                                                            \
    }
        DPCA_I8_SHAPES(X)
#undef X
#undef OCC_W
        // an unknown answer (no device) is not cached
        bool known = false;
        for (int w = 0; w < 2; w++)
            for (int i = 0; i < 3; i++) known |= o.zt[w][i] > 0 || o.za[w][i] > 0;
        if (!known) return o;
        cache[ni][k] = o;
        have[ni][k] = true;
    }
    return cache[ni][k];
}

static int wave_index(int waves) { return waves == 4 ? 1 : 0; }

bool zt_i8_valid(int k, int ns, int sa_tiles, int64_t ldb, int waves)
{
    if ((ns != 2 && ns != 3) || k < 1 || k > kMaxKp || (waves != 8 && waves != 4)) return false;
    const int vt = col_tiles(k, ns).vt();
    return (sa_tiles == 8 || sa_tiles == 4 || sa_tiles == 2) && zt_fits(sa_tiles, vt, waves) &&
           ldb % zt_tile_bytes(sa_tiles, waves) == 0;
}

bool za_i8_valid(int k, int ns, int st_tiles, int waves)
{
    if ((ns != 2 && ns != 3) || k < 1 || k > kMaxKp || (waves != 8 && waves != 4)) return false;
    return (st_tiles == 4 || st_tiles == 2 || st_tiles == 1) &&
           za_fits(st_tiles, col_tiles(k, ns).vt(), waves);
}

// The largest tile count that fits, stepping down while the next smaller
// form doubles the workgroups resident per CU (more tiles per wave reuse
// each LDS operand word more often; more resident waves hide latency).  For
// Z^T T the step down pays only with at most 4 column tiles (measured: k 20
// with 3 slices 1.52 -> 1.23 ms at 20k x 500k; k 40 with 2 slices, 5 tiles,
// 13.2 ms with 4 sample tiles against 14.1 ms with 2 at 100k x 1M).
int zt_i8_tiles(int k, int ns, int64_t ldb, int waves)
{
    const I8Occ o = i8_occ(k, ns);
    const int wi = wave_index(waves);
    const bool narrow = col_tiles(k, ns).vt() <= 4;
    int pick = -1;
    for (int i = 0; i < 3; i++) {
        if (!zt_i8_valid(k, ns, kZtTiles[i], ldb, waves)) continue;
        if (pick < 0 || (narrow && o.zt[wi][pick] > 0 && o.zt[wi][i] >= 2 * o.zt[wi][pick]))
            pick = i;
        else
            break;
    }
    return pick < 0 ? 2 : kZtTiles[pick];
}

int za_i8_tiles(int k, int ns, int waves)
{
    const I8Occ o = i8_occ(k, ns);
    const int wi = wave_index(waves);
    int pick = -1;
    for (int i = 0; i < 3; i++) {
        if (!za_i8_valid(k, ns, kZaTiles[i], waves)) continue;
        if (pick < 0 || (o.za[wi][pick] > 0 && o.za[wi][i] >= 2 * o.za[wi][pick])) pick = i;
        else break;
    }
    return pick < 0 ? 1 : kZaTiles[pick];
}

int zt_i8_blocks_per_cu(int k, int ns, int sa_tiles, int waves)
{
    const I8Occ o = i8_occ(k, ns);
    for (int i = 0; i < 3; i++)
        if (kZtTiles[i] == sa_tiles) return o.zt[wave_index(waves)][i];
    return 0;
}

int64_t zt_i8_tile_bytes(int sa_tiles, int waves) { return zt_tile_bytes(sa_tiles, waves); }

// Tile-order image of rows x ldp packed bytes for the Z^T T workgroups of
// sa_tiles x waves (8 waves only): rows x ldp bytes at img.
hipError_t launch_zt_image(hipStream_t st, const uint8_t *packed, int64_t rows, int64_t ldp,
                           int sa_tiles, int waves, uint8_t *img)
{
    const int tb = zt_tile_bytes(sa_tiles, waves);
    if (waves != 8 || rows % 128 || ldp % tb || !img) return hipErrorInvalidValue;
    const int64_t nblk = rows / 128 * (ldp / tb);
    const unsigned grid = (unsigned)std::min<int64_t>(nblk, 1 << 16);
    uint32_t *out = reinterpret_cast<uint32_t *>(img);
    if (tb == 64) k_zt_image<64><<<grid, 256, 0, st>>>(packed, ldp, nblk, out);
    else if (tb == 128) k_zt_image<128><<<grid, 256, 0, st>>>(packed, ldp, nblk, out);
    else if (tb == 256) k_zt_image<256><<<grid, 256, 0, st>>>(packed, ldp, nblk, out);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

// ldp a multiple of the tile's byte width, rows and rows_per_split
// multiples of 128, rows_per_split at most 2^22 (int32 sums).  img: the
// tile-order image (launch_zt_image, same sa_tiles and waves), or null to
// stage the codes from packed.
hipError_t launch_zt_i8(hipStream_t st, const uint8_t *packed, const uint8_t *img, int64_t rows,
                        int64_t ldp, const void *UW8, const unsigned *smax, int k, int ns,
                        int sa_tiles, int waves, int64_t rows_per_split, int nsplit, float *part)
{
    if (!zt_i8_valid(k, ns, sa_tiles, ldp, waves) || rows % 128 || rows_per_split % 128 ||
        rows_per_split > (1 << 22) || (img && waves != 8))
        return hipErrorInvalidValue;
    const ColTiles ct = col_tiles(k, ns);
    // 1-D grid of tiles x splits, ordered per XCD inside the kernel
    dim3 grid((unsigned)(ldp / zt_tile_bytes(sa_tiles, waves) * nsplit));
    const int64_t Mpad = 4 * ldp;
    const u32x4 *uw = reinterpret_cast<const u32x4 *>(UW8);
#define ZT_T(NS_, SA_, F_, RT_, W_)                                                         \
    if (sa_tiles == SA_ && waves == W_)                                                      \
        return zt_go<NS_, SA_, F_, RT_, W_>(st, grid, packed, img, ldp, uw, smax, rows,      \
                                            rows_per_split, Mpad, ct.R, part);
#define X(NS_, F_, RT_)                                                                     \
    if (ct.NS == NS_ && ct.F == F_ && ct.RT == RT_) {                                        \
        ZT_T(NS_, 8, F_, RT_, 8) ZT_T(NS_, 4, F_, RT_, 8) ZT_T(NS_, 2, F_, RT_, 8)           \
        ZT_T(NS_, 8, F_, RT_, 4) ZT_T(NS_, 4, F_, RT_, 4) ZT_T(NS_, 2, F_, RT_, 4)           \
        return hipErrorInvalidValue;                                                         \
    }
    DPCA_I8_SHAPES(X)
#undef X
#undef ZT_T
    return hipErrorInvalidValue;
}

// rows a multiple of waves*16*st_tiles, seg_bytes of 128, segments of at
// most 65536 samples (int32 sums).
hipError_t launch_za_i8(hipStream_t st, const uint8_t *packed, int64_t rows, int64_t ldp,
                        const void *Q8, const unsigned *qmax, const double *mu,
                        const double *dsc, int k, int ns, int st_tiles, int waves,
                        int64_t seg_bytes, int nseg, float *Tpart)
{
    const int64_t per = (int64_t)waves * 16 * st_tiles;
    // |A| <= 126 per sample: a segment of <= 2^16 samples stays exact in int32
    if (!za_i8_valid(k, ns, st_tiles, waves) || rows % per || seg_bytes % 128 || ldp % 128 ||
        seg_bytes > 16384)
        return hipErrorInvalidValue;
    const ColTiles ct = col_tiles(k, ns);
    // 1-D grid of row blocks x segments, ordered per XCD inside the kernel
    dim3 grid((unsigned)(rows / per * nseg));
    const u32x4 *q = reinterpret_cast<const u32x4 *>(Q8);
    // the last remainder tile holds only the last slice when the other
    // slices of the R remainder columns fit the tiles before it
    const bool skr = ct.RT > 0 && (ns - 1) * ct.R <= 16 * (ct.RT - 1);
#define ZA_T(NS_, ST_, F_, RT_, W_)                                                         \
    if (st_tiles == ST_ && waves == W_) {                                                    \
        if (skr)                                                                             \
            return za_go<NS_, ST_, F_, RT_, true, W_>(st, grid, packed, ldp, q, qmax, mu,    \
                                                      dsc, seg_bytes, rows, ct.R, Tpart);    \
        return za_go<NS_, ST_, F_, RT_, false, W_>(st, grid, packed, ldp, q, qmax, mu, dsc,  \
                                                   seg_bytes, rows, ct.R, Tpart);            \
    }
#define X(NS_, F_, RT_)                                                                     \
    if (ct.NS == NS_ && ct.F == F_ && ct.RT == RT_) {                                        \
        ZA_T(NS_, 4, F_, RT_, 8) ZA_T(NS_, 2, F_, RT_, 8) ZA_T(NS_, 1, F_, RT_, 8)           \
        ZA_T(NS_, 4, F_, RT_, 4) ZA_T(NS_, 2, F_, RT_, 4) ZA_T(NS_, 1, F_, RT_, 4)           \
        return hipErrorInvalidValue;                                                         \
    }
    DPCA_I8_SHAPES(X)
#undef X
#undef ZA_T
    return hipErrorInvalidValue;
}

}  // namespace dpca
