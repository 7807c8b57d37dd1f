// Mixed-precision form of one pass  A A^T Q = Z^T (Z Q)  on bf16 MFMA with
// FP32 accumulation (v_mfma_f32_16x16x32_bf16).
//
//   T = Z Q      A operand: Z = hi + lo, each bf16, looked up from a per-SNP
//                4-entry table (zval split); B operand: Q = hi + lo, bf16.
//                Three products per tile (hi*hi + hi*lo + lo*hi), FP32
//                accumulation over one sample segment; segments summed in
//                FP64.
//   Z^T T        Z_si = d_s (h_si - mu_s m_si) with h the genotype count
//                (0 when missing) and m the non-missing indicator, so
//                Z^T T = sum_s [h_si U_s + m_si W_s],  U = d T, W = -d mu T.
//                The raw codes enter the MFMA unscaled: each SNP takes two
//                K slots (h and m, both exact in bf16) against the bf16 hi+lo
//                split of U and W, two products per tile.  SNP splits are
//                summed in FP64.
//
// MFMA operand layout (16x16x32): lane l = 16g + r holds A[r][8g..8g+7] and
// B[8g..8g+7][r] and receives D[4g + q][r].  Which sample (or SNP) occupies
// which K slot only has to agree between A and B.
#include "dpca_internal.h"
#include "k_common.h"

namespace dpca {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint32_t bf16_bits(float f)  // round to nearest even
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}

// hi = bf16(x), lo = bf16(x - hi) for a double x
__device__ __forceinline__ void split_bf16(double x, uint32_t &hi, uint32_t &lo)
{
    hi = bf16_bits((float)x);
    const double h = (double)__uint_as_float(hi << 16);
    lo = bf16_bits((float)(x - h));
}

__device__ __forceinline__ f4 mfma_bf16(const u32x4 &a, const u32x4 &b, const f4 &c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// Spread the four 2-bit codes of byte b into the four bytes of a dword.
__device__ __forceinline__ uint32_t spread4(uint32_t b)
{
    uint32_t v = (b | (b << 12)) & 0x000F000Fu;
    return (v | (v << 6)) & 0x03030303u;
}

// ----------------------------------------------------------- per-SNP tables --
// ztab[s] = {hi(z0)|hi(z1)<<16, hi(z2)|hi(z3)<<16, lo(z0)|lo(z1)<<16, lo(z2)|lo(z3)<<16}
// where z_c = zval[4s + c] is the reference value for 2-bit code c.
__global__ void k_ztab(const double *__restrict__ zval, int64_t rows, uint4 *__restrict__ ztab)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= rows) return;
    uint32_t h[4], l[4];
#pragma unroll
    for (int c = 0; c < 4; c++) split_bf16(zval[4 * s + c], h[c], l[c]);
    ztab[s] = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), l[0] | (l[1] << 16),
                         l[2] | (l[3] << 16));
}

hipError_t launch_ztab(hipStream_t st, const double *zval, int64_t rows, uint4 *ztab)
{
    k_ztab<<<(unsigned)((rows + 255) / 256), 256, 0, st>>>(zval, rows, ztab);
    return hipGetLastError();
}

// --------------------------------------------------------------- Q split ----
// QB[((i/8)*2 + 0)*kp + c] = 8 bf16 hi of Q[i..i+7][c], [((i/8)*2 + 1)*kp + c]
// = the 8 lo parts: hi and lo rows apart, so 16 lanes reading 16 columns
// touch 256 contiguous bytes of LDS (no bank conflicts).
__global__ void k_qsplit(const double *__restrict__ Q, int64_t groups, int kp,
                         uint4 *__restrict__ QB)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= groups * kp) return;
    const int64_t g8 = idx / kp;
    const int c = (int)(idx % kp);
    uint32_t h[8], l[8];
#pragma unroll
    for (int e = 0; e < 8; e++) split_bf16(Q[(g8 * 8 + e) * kp + c], h[e], l[e]);
    QB[(g8 * 2) * kp + c] = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16),
                                       h[4] | (h[5] << 16), h[6] | (h[7] << 16));
    QB[(g8 * 2 + 1) * kp + c] = make_uint4(l[0] | (l[1] << 16), l[2] | (l[3] << 16),
                                           l[4] | (l[5] << 16), l[6] | (l[7] << 16));
}

hipError_t launch_qsplit(hipStream_t st, const double *Q, int64_t Mpad, int kp, uint4 *QB)
{
    const int64_t n = Mpad / 8 * kp;
    k_qsplit<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(Q, Mpad / 8, kp, QB);
    return hipGetLastError();
}

// --------------------------------------------------------------- T = Z Q ----
// Workgroup: 8 waves x ST tiles x 16 SNPs, one sample segment (blockIdx.y).
// Per 256-sample chunk, lane group g owns the 64 samples of its 16 packed
// bytes; at K step t it supplies samples 64g + 8t .. 64g + 8t + 7.  The bf16
// hi/lo pairs of the chunk's Q rows are staged in LDS.
template <int ST, int CT>
__global__ __launch_bounds__(512) void k_za_mx(const uint8_t *__restrict__ packed, int64_t ldp,
                                               const uint4 *__restrict__ ztab,
                                               const uint4 *__restrict__ QB,
                                               int64_t seg_bytes, int64_t Npad,
                                               float *__restrict__ Tpart)
{
    constexpr int KP = CT * 16;
    __shared__ u32x4 qs[32 * KP * 2];  // [group of 8 samples][hi, lo][col]
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int r = lane & 15, g = lane >> 4;
    const int64_t base = (int64_t)blockIdx.x * (8 * ST * 16) + wave * (ST * 16);
    const int64_t b0 = (int64_t)blockIdx.y * seg_bytes;
    int64_t b1 = b0 + seg_bytes;
    if (b1 > ldp) b1 = ldp;

    u32x4 tab[ST];
    f4 acc[ST][CT];
#pragma unroll
    for (int s = 0; s < ST; s++) {
        tab[s] = reinterpret_cast<const u32x4 *>(ztab)[base + s * 16 + r];
#pragma unroll
        for (int c = 0; c < CT; c++) acc[s][c] = f4{0, 0, 0, 0};
    }

    // Software pipeline: the next chunk's Q operand and packed bytes are
    // loaded into registers while the current chunk is multiplied.
    constexpr int NQ = 32 * KP * 2 / 512;
    static_assert(32 * KP * 2 % 512 == 0, "tile");
    u32x4 rq[NQ], rp[ST];
#define DPCA_ZA_LOAD(c0)                                                                    \
    do {                                                                                    \
        const u32x4 *src_ = reinterpret_cast<const u32x4 *>(QB) + ((c0) / 2) * KP * 2;      \
        _Pragma("unroll") for (int u = 0; u < NQ; u++) rq[u] = src_[tid + 512 * u];        \
        _Pragma("unroll") for (int s = 0; s < ST; s++) rp[s] =                             \
            *reinterpret_cast<const u32x4 *>(packed + (base + s * 16 + r) * ldp + (c0) +     \
                                             16 * g);                                       \
    } while (0)
    DPCA_ZA_LOAD(b0 < b1 ? b0 : 0);

    for (int64_t cb = b0; cb < b1; cb += 64) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NQ; u++) qs[tid + 512 * u] = rq[u];
        u32x4 pk[ST];
#pragma unroll
        for (int s = 0; s < ST; s++) pk[s] = rp[s];
        __syncthreads();
        {
            const int64_t nx = (cb + 64 < b1) ? cb + 64 : cb;
            DPCA_ZA_LOAD(nx);
        }
#pragma unroll 2
        for (int t = 0; t < 8; t++) {
            u32x4 bh[CT], bl[CT];
#pragma unroll
            for (int c = 0; c < CT; c++) {
                bh[c] = qs[((8 * g + t) * 2) * KP + c * 16 + r];
                bl[c] = qs[((8 * g + t) * 2 + 1) * KP + c * 16 + r];
            }
#pragma unroll
            for (int s = 0; s < ST; s++) {
                const uint32_t w = pk[s][t >> 1];
                const uint32_t x16 = (t & 1) ? (w >> 16) : (w & 0xFFFFu);
                const uint32_t c03 = spread4(x16 & 0xFFu), c47 = spread4(x16 >> 8);
                uint32_t sel[4];
                sel[0] = (__builtin_amdgcn_perm(c03, c03, 0x01010000u) << 1) + 0x01000100u;
                sel[1] = (__builtin_amdgcn_perm(c03, c03, 0x03030202u) << 1) + 0x01000100u;
                sel[2] = (__builtin_amdgcn_perm(c47, c47, 0x01010000u) << 1) + 0x01000100u;
                sel[3] = (__builtin_amdgcn_perm(c47, c47, 0x03030202u) << 1) + 0x01000100u;
                u32x4 ah, al;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    ah[j] = __builtin_amdgcn_perm(tab[s][1], tab[s][0], sel[j]);
                    al[j] = __builtin_amdgcn_perm(tab[s][3], tab[s][2], sel[j]);
                }
#pragma unroll
                for (int c = 0; c < CT; c++) {
                    acc[s][c] = mfma_bf16(ah, bh[c], acc[s][c]);
                    acc[s][c] = mfma_bf16(ah, bl[c], acc[s][c]);
                    acc[s][c] = mfma_bf16(al, bh[c], acc[s][c]);
                }
            }
        }
    }
    float *out = Tpart + (int64_t)blockIdx.y * Npad * KP;
#pragma unroll
    for (int s = 0; s < ST; s++)
#pragma unroll
        for (int c = 0; c < CT; c++)
#pragma unroll
            for (int q = 0; q < 4; q++)
                out[(base + s * 16 + 4 * g + q) * KP + c * 16 + r] = acc[s][c][q];
}

template <int ST>
static hipError_t za_mx_ct(hipStream_t st, dim3 grid, const uint8_t *packed, int64_t ldp,
                           const uint4 *ztab, const uint4 *QB, int kp, int64_t seg,
                           int64_t Npad, float *Tpart)
{
    switch (kp) {
    case 16: k_za_mx<ST, 1><<<grid, 512, 0, st>>>(packed, ldp, ztab, QB, seg, Npad, Tpart); break;
    case 32: k_za_mx<ST, 2><<<grid, 512, 0, st>>>(packed, ldp, ztab, QB, seg, Npad, Tpart); break;
    case 48: k_za_mx<ST, 3><<<grid, 512, 0, st>>>(packed, ldp, ztab, QB, seg, Npad, Tpart); break;
    case 64: k_za_mx<ST, 4><<<grid, 512, 0, st>>>(packed, ldp, ztab, QB, seg, Npad, Tpart); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// rows must be a multiple of 8*16*st_tiles; seg_bytes a multiple of 64.
hipError_t launch_za_mx(hipStream_t st, const uint8_t *packed, int64_t rows, int64_t ldp,
                        const uint4 *ztab, const uint4 *QB, int kp, int st_tiles,
                        int64_t seg_bytes, int nseg, float *Tpart)
{
    if (seg_bytes % 64 || ldp % 64) return hipErrorInvalidValue;
    const int64_t per = 8 * 16 * st_tiles;
    if (rows % per) return hipErrorInvalidValue;
    dim3 grid((unsigned)(rows / per), (unsigned)nseg);
    if (st_tiles == 8) return za_mx_ct<8>(st, grid, packed, ldp, ztab, QB, kp, seg_bytes, rows, Tpart);
    if (st_tiles == 4) return za_mx_ct<4>(st, grid, packed, ldp, ztab, QB, kp, seg_bytes, rows, Tpart);
    if (st_tiles == 2) return za_mx_ct<2>(st, grid, packed, ldp, ztab, QB, kp, seg_bytes, rows, Tpart);
    return hipErrorInvalidValue;
}

// ------------------------------------------------ T segments -> T, U, W -----
// T[s][c] = sum of the segment partials (FP64); U = d_s T, W = -d_s mu_s T;
// UWB[((s/4)*2 + 0)*kp + c] = bf16 hi of U(s0..s3), W(s0..s3);
// UWB[((s/4)*2 + 1)*kp + c] = the lo parts.
__global__ void k_tprep(const float *__restrict__ Tpart, int nseg, int64_t rows, int kp,
                        const double *__restrict__ mu, const double *__restrict__ dsc,
                        double *__restrict__ T, uint4 *__restrict__ UWB)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= rows / 4 * kp) return;
    const int64_t s4 = idx / kp;
    const int c = (int)(idx % kp);
    uint32_t uh[4], ul[4], wh[4], wl[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int64_t s = s4 * 4 + e;
        double t = 0;
        for (int p = 0; p < nseg; p++) t += (double)Tpart[((int64_t)p * rows + s) * kp + c];
        T[s * kp + c] = t;
        const double u = dsc[s] * t;
        const double w = -(dsc[s] * mu[s]) * t;
        split_bf16(u, uh[e], ul[e]);
        split_bf16(w, wh[e], wl[e]);
    }
    UWB[(s4 * 2) * kp + c] = make_uint4(uh[0] | (uh[1] << 16), uh[2] | (uh[3] << 16),
                                        wh[0] | (wh[1] << 16), wh[2] | (wh[3] << 16));
    UWB[(s4 * 2 + 1) * kp + c] = make_uint4(ul[0] | (ul[1] << 16), ul[2] | (ul[3] << 16),
                                            wl[0] | (wl[1] << 16), wl[2] | (wl[3] << 16));
}

hipError_t launch_tprep(hipStream_t st, const float *Tpart, int nseg, int64_t rows, int kp,
                        const double *mu, const double *dsc, double *T, uint4 *UWB)
{
    const int64_t n = rows / 4 * kp;
    k_tprep<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(Tpart, nseg, rows, kp, mu, dsc, T, UWB);
    return hipGetLastError();
}

// U/W operand of a resident FP64 right-hand matrix (projection loadings):
// same layout as k_tprep's output.
__global__ void k_uwb(const double *__restrict__ T, int64_t rows, int kp,
                      const double *__restrict__ mu, const double *__restrict__ dsc,
                      uint4 *__restrict__ UWB)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= rows / 4 * kp) return;
    const int64_t s4 = idx / kp;
    const int c = (int)(idx % kp);
    uint32_t uh[4], ul[4], wh[4], wl[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int64_t s = s4 * 4 + e;
        const double t = T[s * kp + c];
        split_bf16(dsc[s] * t, uh[e], ul[e]);
        split_bf16(-(dsc[s] * mu[s]) * t, wh[e], wl[e]);
    }
    UWB[(s4 * 2) * kp + c] = make_uint4(uh[0] | (uh[1] << 16), uh[2] | (uh[3] << 16),
                                        wh[0] | (wh[1] << 16), wh[2] | (wh[3] << 16));
    UWB[(s4 * 2 + 1) * kp + c] = make_uint4(ul[0] | (ul[1] << 16), ul[2] | (ul[3] << 16),
                                            wl[0] | (wl[1] << 16), wl[2] | (wl[3] << 16));
}

hipError_t launch_uwb(hipStream_t st, const double *T, int64_t rows, int kp, const double *mu,
                      const double *dsc, uint4 *UWB)
{
    const int64_t n = rows / 4 * kp;
    k_uwb<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(T, rows, kp, mu, dsc, UWB);
    return hipGetLastError();
}

// ------------------------------------------------------------ Z^T T part ----
// Workgroup: 8 waves x SA tiles x 16 samples (SA = 8: 1024 samples = 256
// packed bytes per row) and one SNP split (blockIdx.y).  128-SNP chunks of
// packed bytes are staged in LDS transposed to [byte column][SNP], so a lane
// reads the bytes of 4 consecutive SNPs for its sample with one dword load.
// K step t (16 SNPs): lane group g holds SNPs 16t + 4g .. +3; K slots 0..3 are
// their h, slots 4..7 their m.
template <int SA, int CT>
__global__ __launch_bounds__(512) void k_zt_mx(const uint8_t *__restrict__ packed, int64_t ldp,
                                               const uint4 *__restrict__ UWB, int64_t rows,
                                               int64_t rows_per_split, int64_t Mpad,
                                               float *__restrict__ part)
{
    constexpr int KP = CT * 16;
    constexpr int TB = SA * 16 * 8 / 4;    // packed bytes per row in this tile
    constexpr int SC = 128;                // SNPs per chunk
    // [byte column][4-SNP group], rows padded by one dword so that the
    // transposing stores (consecutive byte columns) and the reads (4 byte
    // columns x 4 groups per 16 lanes) spread over the LDS banks
    __shared__ uint32_t pk[TB][SC / 4 + 1];
    __shared__ u32x4 uw[(SC / 4) * KP * 2];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int r = lane & 15, g = lane >> 4;
    const int64_t mb = (int64_t)blockIdx.x * TB;
    const int64_t sbeg = (int64_t)blockIdx.y * rows_per_split;
    int64_t send = sbeg + rows_per_split;
    if (send > rows) send = rows;
    // h / m tables for perm: bytes indexed by the 2-bit code
    const uint32_t HL = 0x00800000u, HH = 0x403F0000u, ML = 0x80800080u, MH = 0x3F3F003Fu;

    f4 acc[SA][CT];
#pragma unroll
    for (int a = 0; a < SA; a++)
#pragma unroll
        for (int c = 0; c < CT; c++) acc[a][c] = f4{0, 0, 0, 0};

    // Software pipeline: the next chunk's packed rows and U/W operands are
    // loaded into registers while the current chunk is multiplied out of LDS.
    constexpr int NPK = (SC / 4) * (TB / 4) / 512;   // 4x4 byte blocks per thread
    constexpr int NUW = (SC / 4) * KP * 2 / 512;     // uint4 of U/W per thread
    static_assert((SC / 4) * (TB / 4) % 512 == 0 && (SC / 4) * KP * 2 % 512 == 0, "tile");
    uint32_t rpk[NPK][4];
    u32x4 ruw[NUW];
#define DPCA_ZT_LOAD(c0)                                                                    \
    do {                                                                                    \
        _Pragma("unroll") for (int u = 0; u < NPK; u++) {                                  \
            const int it_ = tid + 512 * u;                                                  \
            const int sg_ = it_ / (TB / 4), bq_ = it_ % (TB / 4);                           \
            const uint8_t *p_ = packed + ((c0) + 4 * sg_) * ldp + mb + 4 * bq_;             \
            _Pragma("unroll") for (int j = 0; j < 4; j++) rpk[u][j] =                      \
                *reinterpret_cast<const uint32_t *>(p_ + j * ldp);                          \
        }                                                                                   \
        const u32x4 *q_ = reinterpret_cast<const u32x4 *>(UWB) + ((c0) / 4) * KP * 2;      \
        _Pragma("unroll") for (int u = 0; u < NUW; u++) ruw[u] = q_[tid + 512 * u];        \
    } while (0)
    DPCA_ZT_LOAD(sbeg < send ? sbeg : 0);

    for (int64_t s0 = sbeg; s0 < send; s0 += SC) {
        __syncthreads();
        // 4x4 byte transpose of each block into [byte column][SNP group]
#pragma unroll
        for (int u = 0; u < NPK; u++) {
            const int it = tid + 512 * u;
            const int sg = it / (TB / 4), bq = it % (TB / 4);
            const uint32_t r0 = rpk[u][0], r1 = rpk[u][1], r2 = rpk[u][2], r3 = rpk[u][3];
            const uint32_t lo01 = __builtin_amdgcn_perm(r1, r0, 0x05010400u);
            const uint32_t hi01 = __builtin_amdgcn_perm(r1, r0, 0x07030602u);
            const uint32_t lo23 = __builtin_amdgcn_perm(r3, r2, 0x05010400u);
            const uint32_t hi23 = __builtin_amdgcn_perm(r3, r2, 0x07030602u);
            pk[4 * bq + 0][sg] = __builtin_amdgcn_perm(lo23, lo01, 0x05040100u);
            pk[4 * bq + 1][sg] = __builtin_amdgcn_perm(lo23, lo01, 0x07060302u);
            pk[4 * bq + 2][sg] = __builtin_amdgcn_perm(hi23, hi01, 0x05040100u);
            pk[4 * bq + 3][sg] = __builtin_amdgcn_perm(hi23, hi01, 0x07060302u);
        }
#pragma unroll
        for (int u = 0; u < NUW; u++) uw[tid + 512 * u] = ruw[u];
        __syncthreads();
        {
            // unconditional (re-reads the last chunk at the end) so the
            // prefetch registers stay registers
            const int64_t nx = (s0 + SC < send) ? s0 + SC : s0;
            DPCA_ZT_LOAD(nx);
        }
#pragma unroll 1
        for (int t = 0; t < SC / 16; t++) {
            u32x4 bh[CT], bl[CT];
#pragma unroll
            for (int c = 0; c < CT; c++) {
                bh[c] = uw[((4 * t + g) * 2) * KP + c * 16 + r];
                bl[c] = uw[((4 * t + g) * 2 + 1) * KP + c * 16 + r];
            }
#pragma unroll
            for (int a = 0; a < SA; a++) {
                const int i = wave * SA * 16 + a * 16 + r;  // sample within the tile
                const uint32_t wv = pk[i >> 2][4 * t + g];
                const uint32_t cw = (wv >> (2 * (i & 3))) & 0x03030303u;
                const uint32_t s01 = __builtin_amdgcn_perm(cw, cw, 0x01010000u) + 0x04000400u;
                const uint32_t s23 = __builtin_amdgcn_perm(cw, cw, 0x03030202u) + 0x04000400u;
                const u32x4 av = u32x4{__builtin_amdgcn_perm(HH, HL, s01),
                                       __builtin_amdgcn_perm(HH, HL, s23),
                                       __builtin_amdgcn_perm(MH, ML, s01),
                                       __builtin_amdgcn_perm(MH, ML, s23)};
#pragma unroll
                for (int c = 0; c < CT; c++) {
                    acc[a][c] = mfma_bf16(av, bh[c], acc[a][c]);
                    acc[a][c] = mfma_bf16(av, bl[c], acc[a][c]);
                }
            }
        }
    }
    float *out = part + ((int64_t)blockIdx.y * Mpad + 4 * mb + wave * SA * 16) * KP;
#pragma unroll
    for (int a = 0; a < SA; a++)
#pragma unroll
        for (int c = 0; c < CT; c++)
#pragma unroll
            for (int q = 0; q < 4; q++) out[(a * 16 + 4 * g + q) * KP + c * 16 + r] = acc[a][c][q];
}

template <int SA>
static hipError_t zt_mx_ct(hipStream_t st, dim3 grid, const uint8_t *packed, int64_t ldp,
                           const uint4 *UWB, int kp, int64_t rows, int64_t rps, int64_t Mpad,
                           float *part)
{
    switch (kp) {
    case 16: k_zt_mx<SA, 1><<<grid, 512, 0, st>>>(packed, ldp, UWB, rows, rps, Mpad, part); break;
    case 32: k_zt_mx<SA, 2><<<grid, 512, 0, st>>>(packed, ldp, UWB, rows, rps, Mpad, part); break;
    case 48: k_zt_mx<SA, 3><<<grid, 512, 0, st>>>(packed, ldp, UWB, rows, rps, Mpad, part); break;
    case 64: k_zt_mx<SA, 4><<<grid, 512, 0, st>>>(packed, ldp, UWB, rows, rps, Mpad, part); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ldp must be a multiple of the tile's byte width (SA*32), rows of 128,
// rows_per_split a multiple of 128.
hipError_t launch_zt_mx(hipStream_t st, const uint8_t *packed, int64_t rows, int64_t ldp,
                        const uint4 *UWB, int kp, int sa_tiles, int64_t rows_per_split,
                        int nsplit, float *part)
{
    const int64_t tb = (int64_t)sa_tiles * 32;
    if (ldp % tb || rows % 128 || rows_per_split % 128) return hipErrorInvalidValue;
    dim3 grid((unsigned)(ldp / tb), (unsigned)nsplit);
    const int64_t Mpad = 4 * ldp;
    if (sa_tiles == 8) return zt_mx_ct<8>(st, grid, packed, ldp, UWB, kp, rows, rows_per_split, Mpad, part);
    if (sa_tiles == 2) return zt_mx_ct<2>(st, grid, packed, ldp, UWB, kp, rows, rows_per_split, Mpad, part);
    return hipErrorInvalidValue;
}

// out[e] (+)= sum over splits of float partials, in FP64, split order.
__global__ void k_reduce_f32(const float *__restrict__ part, int splits, int64_t n,
                             double *__restrict__ out, int accumulate)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double s = accumulate ? out[e] : 0.0;
    for (int p = 0; p < splits; p++) s += (double)part[(int64_t)p * n + e];
    out[e] = s;
}

hipError_t launch_reduce_f32(hipStream_t st, const float *part, int splits, int64_t n,
                             double *out, int accumulate)
{
    k_reduce_f32<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(part, splits, n, out, accumulate);
    return hipGetLastError();
}

}  // namespace dpca
