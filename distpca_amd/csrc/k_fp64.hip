// Packed-genotype kernels of libdpca: repack into the padded HBM layout,
// per-SNP statistics with the reference's value table, decode, and the FP64
// fused decode + MFMA products of one pass  A A^T Q = Z^T (Z Q).
//
// Z is never materialised: each kernel reads the 2-bit codes and turns them
// into Z values through the per-SNP 4-entry table zval (exactly the values the
// reference's decode + standardise produce, src/methods.cpp:814-845).
#include "dpca_internal.h"
#include "k_common.h"

#include <atomic>

namespace dpca {

// ------------------------------------------------------------------ repack --
// raw rows of np bytes (the .bed layout after the 3-byte magic) -> rows of ldp
// bytes.  Bytes past np are 0x55 (four missing codes); within the last real
// byte, the codes of samples >= M are forced to 01 (missing) so that padding
// never enters any sum.  One thread per destination byte.
__global__ void k_repack(const uint8_t *__restrict__ raw, int64_t rows, int64_t np,
                         uint8_t *__restrict__ dst, int64_t ldp, int M)
{
    const int64_t total = rows * ldp;
    DPCA_EW_LOOP(idx, total) {
        const int64_t r = idx / ldp, b = idx - r * ldp;
        uint8_t v = kPadByte;
        if (b < np) {
            v = raw[r * np + b];
            if (b == np - 1) {
                const int valid = M - 4 * (int)(np - 1);  // 1..4 real samples in this byte
                if (valid < 4) {
                    const uint8_t keep = (uint8_t)((1u << (2 * valid)) - 1u);
                    v = (uint8_t)((v & keep) | (kPadByte & ~keep));
                }
            }
        }
        dst[idx] = v;
    }
}

__global__ void k_fill_u8(uint8_t *p, int64_t n, uint8_t v)
{
    DPCA_EW_LOOP(i, n) p[i] = v;
}

__global__ void k_fill_f64(double *p, int64_t n, double v)
{
    DPCA_EW_LOOP(i, n) p[i] = v;
}

hipError_t launch_repack(hipStream_t st, const uint8_t *raw, int64_t rows, int64_t np,
                         uint8_t *dst, int64_t ldp, int M)
{
    int64_t total = rows * ldp;
    if (total == 0) return hipSuccess;
    k_repack<<<ew_blocks(total), 256, 0, st>>>(raw, rows, np, dst, ldp, M);
    return hipGetLastError();
}

hipError_t launch_fill_u8(hipStream_t st, uint8_t *p, int64_t n, uint8_t v)
{
    if (n == 0) return hipSuccess;
    k_fill_u8<<<ew_blocks(n), 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_f64(hipStream_t st, double *p, int64_t n, double v)
{
    if (n == 0) return hipSuccess;
    k_fill_f64<<<ew_blocks(n), 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}

// M x k (dense) <-> Mpad x kp (zero padded) for the right-hand sides.
__global__ void k_pad_rhs(const double *__restrict__ src, int M, int k, int64_t Mpad, int kp,
                          double *__restrict__ dst)
{
    DPCA_EW_LOOP(e, Mpad * kp) {
        const int64_t i = e / kp;
        const int j = (int)(e % kp);
        dst[e] = (i < M && j < k) ? src[i * k + j] : 0.0;
    }
}

__global__ void k_compact_rhs(const double *__restrict__ src, int M, int k, int kp,
                              double *__restrict__ dst)
{
    DPCA_EW_LOOP(e, (int64_t)M * k) dst[e] = src[(e / k) * kp + e % k];
}

hipError_t launch_pad_rhs(hipStream_t st, const double *src, int M, int k, int64_t Mpad, int kp,
                          double *dst)
{
    k_pad_rhs<<<ew_blocks(Mpad * kp), 256, 0, st>>>(src, M, k, Mpad, kp, dst);
    return hipGetLastError();
}

hipError_t launch_compact_rhs(hipStream_t st, const double *src, int M, int k, int kp,
                              double *dst)
{
    k_compact_rhs<<<ew_blocks((int64_t)M * k), 256, 0, st>>>(src, M, k, kp, dst);
    return hipGetLastError();
}

// Projection batches: copy a request's SNP-major payload (rows of nb bytes)
// into byte columns [off, off+nb) of the batch matrix (row pitch ldb);
// sample codes past n_valid in the last byte become 01 (missing).
__global__ void k_scatter_cols(const uint8_t *__restrict__ src, int64_t rows, int64_t nb,
                               int n_valid, uint8_t *__restrict__ dst, int64_t ldb, int64_t off)
{
    DPCA_EW_LOOP(idx, rows * nb) {
        const int64_t r = idx / nb, b = idx - r * nb;
        uint8_t v = src[idx];
        if (b == nb - 1) {
            const int valid = n_valid - 4 * (int)(nb - 1);
            if (valid < 4) {
                const uint8_t keep = (uint8_t)((1u << (2 * valid)) - 1u);
                v = (uint8_t)((v & keep) | (kPadByte & ~keep));
            }
        }
        dst[r * ldb + off + b] = v;
    }
}

// Fill byte columns [c0, c1) of every row with 0x55 (missing codes).
__global__ void k_fill_cols(uint8_t *__restrict__ dst, int64_t rows, int64_t ldb, int64_t c0,
                            int64_t c1)
{
    const int64_t w = c1 - c0;
    DPCA_EW_LOOP(idx, rows * w) {
        const int64_t r = idx / w;
        dst[r * ldb + c0 + (idx - r * w)] = kPadByte;
    }
}

hipError_t launch_scatter_cols(hipStream_t st, const uint8_t *src, int64_t rows, int64_t nb,
                               int n_valid, uint8_t *dst, int64_t ldb, int64_t off)
{
    const int64_t n = rows * nb;
    if (n == 0) return hipSuccess;
    k_scatter_cols<<<ew_blocks(n), 256, 0, st>>>(src, rows, nb, n_valid, dst, ldb,
                                                                off);
    return hipGetLastError();
}

hipError_t launch_fill_cols(hipStream_t st, uint8_t *dst, int64_t rows, int64_t ldb, int64_t c0,
                            int64_t c1)
{
    const int64_t n = rows * (c1 - c0);
    if (n <= 0) return hipSuccess;
    k_fill_cols<<<ew_blocks(n), 256, 0, st>>>(dst, rows, ldb, c0, c1);
    return hipGetLastError();
}

// Loadings W = T diag(1/sigma^2) for the first nsv columns, 0 elsewhere.
__global__ void k_loadings(const double *__restrict__ T, int64_t rows, int kp, int nsv,
                           const double *__restrict__ inv_s2, double *__restrict__ W)
{
    DPCA_EW_LOOP(e, rows * kp) {
        const int j = (int)(e % kp);
        W[e] = (j < nsv) ? T[e] * inv_s2[j] : 0.0;
    }
}

hipError_t launch_loadings(hipStream_t st, const double *T, int64_t rows, int kp, int nsv,
                           const double *inv_s2, double *W)
{
    const int64_t n = rows * kp;
    k_loadings<<<ew_blocks(n), 256, 0, st>>>(T, rows, kp, nsv, inv_s2, W);
    return hipGetLastError();
}

// --------------------------------------------------------------- SNP stats --
// One wavefront per SNP row.  Codes are counted 16 at a time with popcounts:
// with lo = bit0 and hi = bit1 of every 2-bit field, missing (01) = lo & ~hi,
// and sum g = popc(hi) + popc(hi & lo) since 10 -> 1 and 11 -> 2.  Integer
// counts are exact; the table is then formed in the reference's own IEEE
// operation order (src/utilities.cpp:288-301, src/methods.cpp:824-830, with
// the * inv_sqrtN of src/methods.cpp:842 folded in).
__global__ __launch_bounds__(256) void k_snp_stats(const uint8_t *__restrict__ packed,
                                                   int64_t rows, int64_t ldp,
                                                   double inv_sqrtN, int64_t *sum,
                                                   int64_t *ctr, double *zval, double *mu,
                                                   double *dsc)
{
#pragma clang fp contract(off)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t s = (int64_t)blockIdx.x * 4 + wave;
    if (s >= rows) return;
    const uint4 *row = reinterpret_cast<const uint4 *>(packed + s * ldp);
    const int nq = (int)(ldp >> 4);
    long long sm = 0, c = 0;
    for (int q = lane; q < nq; q += 64) {
        uint4 v = row[q];
        uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int e = 0; e < 4; e++) {
            uint32_t lo = w[e] & 0x55555555u, hi = (w[e] >> 1) & 0x55555555u;
            uint32_t miss = lo & ~hi;
            c += __popc(0x55555555u & ~miss);
            sm += __popc(hi) + __popc(hi & lo);
        }
    }
    sm = wave_sum_ll(sm);
    c = wave_sum_ll(c);
    if (lane == 0) {
        double a = (double)sm / (double)(unsigned int)c;
        double P = a / 2.0;
        double d = sqrt(2.0 * P * (1 - P));
        sum[s] = sm;
        ctr[s] = c;
        double z0 = 0, z2 = 0, z3 = 0, m = 0, sc = 0;
        if (d > 1e-9) {
            z0 = (0 - a) / d * inv_sqrtN;
            z2 = (1 - a) / d * inv_sqrtN;
            z3 = (2 - a) / d * inv_sqrtN;
            m = a;
            sc = inv_sqrtN / d;
        }
        zval[4 * s + 0] = z0;
        zval[4 * s + 1] = 0;
        zval[4 * s + 2] = z2;
        zval[4 * s + 3] = z3;
        mu[s] = m;
        dsc[s] = sc;
    }
}

hipError_t launch_snp_stats(hipStream_t st, const uint8_t *packed, int64_t rows, int64_t ldp,
                            double inv_sqrtN, int64_t *sum, int64_t *ctr, double *zval,
                            double *mu, double *dsc)
{
    if (rows == 0) return hipSuccess;
    // one wave per row: 64 rows' work-items fit a dispatch up to 67M rows
    if ((rows + 3) / 4 * 256 > (int64_t)UINT32_MAX) return hipErrorInvalidValue;
    k_snp_stats<<<(unsigned)((rows + 3) / 4), 256, 0, st>>>(packed, rows, ldp, inv_sqrtN, sum,
                                                            ctr, zval, mu, dsc);
    return hipGetLastError();
}

// ------------------------------------------------------------------ decode --
// codes[s*M + i] = {0,3,1,2}[2-bit code]  (src/utilities.cpp:353).
__global__ void k_decode_codes(const uint8_t *__restrict__ packed, int64_t rows, int64_t ldp,
                               int M, uint8_t *__restrict__ codes)
{
    const int64_t nb = (M + 3) / 4;
    const uint8_t map[4] = {0, 3, 1, 2};
    DPCA_EW_LOOP(idx, rows * nb) {
        const int64_t s = idx / nb, b = idx - s * nb;
        const uint8_t v = packed[s * ldp + b];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int64_t i = 4 * b + j;
            if (i < M) codes[s * M + i] = map[(v >> (2 * j)) & 3];
        }
    }
}

hipError_t launch_decode_codes(hipStream_t st, const uint8_t *packed, int64_t rows,
                               int64_t ldp, int M, uint8_t *codes)
{
    int64_t n = rows * ((M + 3) / 4);
    if (n == 0) return hipSuccess;
    k_decode_codes<<<ew_blocks(n), 256, 0, st>>>(packed, rows, ldp, M, codes);
    return hipGetLastError();
}

// Dense standardised matrix for the full-SVD diagnostic: out[s * M + i] =
// zval[4 s + code(s, i)] for rows s < rows and samples i < M (the values
// src/utilities.cpp:272-306 produce; SNP-major, i.e. column-major M x rows).
__global__ void k_decode_f64(const uint8_t *__restrict__ packed, int64_t rows, int64_t ldp, int M,
                             const double *__restrict__ zval, double *__restrict__ out)
{
    DPCA_EW_LOOP(e, rows * M) {
        const int64_t s = e / M;
        const int i = (int)(e % M);
        const int code = (packed[s * ldp + (i >> 2)] >> (2 * (i & 3))) & 3;
        out[e] = zval[4 * s + code];
    }
}

hipError_t launch_decode_f64(hipStream_t st, const uint8_t *packed, int64_t rows, int64_t ldp,
                             int M, const double *zval, double *out)
{
    const int64_t n = rows * M;
    if (n == 0) return hipSuccess;
    k_decode_f64<<<ew_blocks(n), 256, 0, st>>>(packed, rows, ldp, M, zval, out);
    return hipGetLastError();
}

// ---------------------------------------------------------- FP64  T = Z Q --
// v_mfma_f64_16x16x4f64: D(16x16) += A(16x4) B(4x16).  Lane l = 16*g + r
// supplies A[r][g] and B[g][r] and holds D[4q + g][r], q = 0..3 (rows
// interleaved by lane group; checked on the device).
// Rows of the tile are SNPs, columns are right-hand sides, and the 4-deep K
// slot g is a sample.  Which sample sits in which K slot is free as long as
// A and B agree: lane group g walks the 64 samples of its own 16 packed
// bytes, so every lane decodes from one 16-byte load per 256-sample chunk.
// A workgroup owns 4 waves x 2 tiles of 16 SNPs (k_za_f64w: x 4 tiles, used
// with two or more whole column tiles where it keeps two workgroups per CU:
// each Q load then feeds twice the MFMAs; 5% faster at nrhs 40, 9% slower at
// nrhs 20) and the whole sample range; Q (Mpad x kp, zero rows past M) is
// read through L1/L2.
// blockIdx.y takes the packed bytes [y seg_bytes, (y + 1) seg_bytes) of every
// row (a sample segment; one segment when the rows alone give enough
// workgroups) and writes its partial T at T + y rows kp.
//
// Columns: CT whole 16-column tiles, then NR groups of 4 columns on
// v_mfma_f64_4x4x4f64 (4 blocks of D(4x4) += A(4x4) B(4x4); lane
// 16k + 4m + i holds A_m[i][k], lane 16k + 4m + j holds B_m[k][j], lane
// 16i + 4m + j receives D_m[i][j]; checked on the device).  With block m =
// row group r / 4 the A operand is the same lane value as in the 16x16 tile
// (row r, K slot g), every block takes the same 4 columns of Q, and a 4x4x4
// costs 17 cycles against 64 for a 16x16x4 whose 12 extra columns would be
// zero padding (nrhs 20 = 16 + 4, 40 = 2 x 16 + 8).  The padding columns
// past the last group are written as zeros.
template <int CT, int NR>
__global__ __launch_bounds__(256) void k_za_f64(const uint8_t *__restrict__ packed,
                                                int64_t ldp, const double *__restrict__ zval,
                                                const double *__restrict__ Q,
                                                double *__restrict__ T, int64_t seg_bytes,
                                                int64_t rows)
{
    constexpr int KP = (CT + (NR > 0 ? 1 : 0)) * 16;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = lane & 15, g = lane >> 4;
    const int64_t base = (int64_t)blockIdx.x * 128 + wave * 32;
    const int64_t s0 = base + r, s1 = base + 16 + r;
    const double a00 = zval[4 * s0 + 0], a01 = zval[4 * s0 + 1], a02 = zval[4 * s0 + 2],
                 a03 = zval[4 * s0 + 3];
    const double a10 = zval[4 * s1 + 0], a11 = zval[4 * s1 + 1], a12 = zval[4 * s1 + 2],
                 a13 = zval[4 * s1 + 3];
    d4 acc0[CT > 0 ? CT : 1], acc1[CT > 0 ? CT : 1];
    double rem0[NR > 0 ? NR : 1], rem1[NR > 0 ? NR : 1];
#pragma unroll
    for (int c = 0; c < CT; c++) {
        acc0[c] = d4{0, 0, 0, 0};
        acc1[c] = d4{0, 0, 0, 0};
    }
#pragma unroll
    for (int n = 0; n < NR; n++) {
        rem0[n] = 0.0;
        rem1[n] = 0.0;
    }
    const uint8_t *p0 = packed + s0 * ldp + 16 * g;
    const uint8_t *p1 = packed + s1 * ldp + 16 * g;
    const int64_t cb0 = (int64_t)blockIdx.y * seg_bytes;
    const int64_t cb1 = cb0 + seg_bytes < ldp ? cb0 + seg_bytes : ldp;
    T += (int64_t)blockIdx.y * rows * KP;
    for (int64_t cb = cb0; cb < cb1; cb += 64) {
        const uint4 v0 = *reinterpret_cast<const uint4 *>(p0 + cb);
        const uint4 v1 = *reinterpret_cast<const uint4 *>(p1 + cb);
        const uint32_t w0[4] = {v0.x, v0.y, v0.z, v0.w};
        const uint32_t w1[4] = {v1.x, v1.y, v1.z, v1.w};
        const double *qrow = Q + (4 * cb + 64 * g) * KP + r;
        const double *qrem = Q + (4 * cb + 64 * g) * KP + CT * 16 + (r & 3);
#pragma unroll 8
        for (int t = 0; t < 64; t++) {
            double b[CT > 0 ? CT : 1], br[NR > 0 ? NR : 1];
#pragma unroll
            for (int c = 0; c < CT; c++) b[c] = qrow[t * KP + c * 16];
#pragma unroll
            for (int n = 0; n < NR; n++) br[n] = qrem[t * KP + 4 * n];
            const uint32_t c0 = code_at(w0, t), c1 = code_at(w1, t);
            const double x0 = sel4(c0, a00, a01, a02, a03);
            const double x1 = sel4(c1, a10, a11, a12, a13);
#pragma unroll
            for (int c = 0; c < CT; c++) {
                acc0[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(x0, b[c], acc0[c], 0, 0, 0);
                acc1[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(x1, b[c], acc1[c], 0, 0, 0);
            }
#pragma unroll
            for (int n = 0; n < NR; n++) {
                rem0[n] = __builtin_amdgcn_mfma_f64_4x4x4f64(x0, br[n], rem0[n], 0, 0, 0);
                rem1[n] = __builtin_amdgcn_mfma_f64_4x4x4f64(x1, br[n], rem1[n], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < CT; c++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            T[(base + 4 * q + g) * KP + c * 16 + r] = acc0[c][q];
            T[(base + 16 + 4 * q + g) * KP + c * 16 + r] = acc1[c][q];
        }
    if constexpr (NR > 0) {
        // D_m[i][j] of this lane: row 4m + i = 4 (r / 4) + g, column r % 4
        const int row = 4 * (r >> 2) + g;
#pragma unroll
        for (int n = 0; n < NR; n++) {
            T[(base + row) * KP + CT * 16 + 4 * n + (r & 3)] = rem0[n];
            T[(base + 16 + row) * KP + CT * 16 + 4 * n + (r & 3)] = rem1[n];
        }
        constexpr int PADC = KP - CT * 16 - 4 * NR;  // zero columns of the last tile
        if constexpr (PADC > 0)
            for (int e = lane; e < 32 * PADC; e += 64)
                T[(base + e / PADC) * KP + CT * 16 + 4 * NR + e % PADC] = 0.0;
    }
}

// The same with ST SNP tiles per wave (used with ST = 4).
template <int CT, int NR, int ST>
__global__ __launch_bounds__(256) void k_za_f64w(const uint8_t *__restrict__ packed,
                                                int64_t ldp, const double *__restrict__ zval,
                                                const double *__restrict__ Q,
                                                double *__restrict__ T, int64_t seg_bytes,
                                                int64_t rows)
{
    constexpr int KP = (CT + (NR > 0 ? 1 : 0)) * 16;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = lane & 15, g = lane >> 4;
    const int64_t base = (int64_t)blockIdx.x * (64 * ST) + wave * (16 * ST);
    double av[ST][4];  // the 4-entry value table of this lane's SNP in each tile
    const uint8_t *pp[ST];
#pragma unroll
    for (int s = 0; s < ST; s++) {
        const int64_t row = base + 16 * s + r;
#pragma unroll
        for (int j = 0; j < 4; j++) av[s][j] = zval[4 * row + j];
        pp[s] = packed + row * ldp + 16 * g;
    }
    d4 acc[ST][CT > 0 ? CT : 1];
    double rem[ST][NR > 0 ? NR : 1];
#pragma unroll
    for (int s = 0; s < ST; s++) {
#pragma unroll
        for (int c = 0; c < CT; c++) acc[s][c] = d4{0, 0, 0, 0};
#pragma unroll
        for (int n = 0; n < NR; n++) rem[s][n] = 0.0;
    }
    const int64_t cb0 = (int64_t)blockIdx.y * seg_bytes;
    const int64_t cb1 = cb0 + seg_bytes < ldp ? cb0 + seg_bytes : ldp;
    T += (int64_t)blockIdx.y * rows * KP;
    for (int64_t cb = cb0; cb < cb1; cb += 64) {
        uint32_t w[ST][4];
#pragma unroll
        for (int s = 0; s < ST; s++) {
            const uint4 v = *reinterpret_cast<const uint4 *>(pp[s] + cb);
            w[s][0] = v.x;
            w[s][1] = v.y;
            w[s][2] = v.z;
            w[s][3] = v.w;
        }
        const double *qrow = Q + (4 * cb + 64 * g) * KP + r;
        const double *qrem = Q + (4 * cb + 64 * g) * KP + CT * 16 + (r & 3);
        constexpr int UNR = ST == 2 ? 8 : 4;
#pragma unroll UNR
        for (int t = 0; t < 64; t++) {
            double b[CT > 0 ? CT : 1], br[NR > 0 ? NR : 1];
#pragma unroll
            for (int c = 0; c < CT; c++) b[c] = qrow[t * KP + c * 16];
#pragma unroll
            for (int n = 0; n < NR; n++) br[n] = qrem[t * KP + 4 * n];
            double x[ST];
#pragma unroll
            for (int s = 0; s < ST; s++)
                x[s] = sel4(code_at(w[s], t), av[s][0], av[s][1], av[s][2], av[s][3]);
#pragma unroll
            for (int c = 0; c < CT; c++)
#pragma unroll
                for (int s = 0; s < ST; s++)
                    acc[s][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[s], b[c], acc[s][c], 0, 0, 0);
#pragma unroll
            for (int n = 0; n < NR; n++)
#pragma unroll
                for (int s = 0; s < ST; s++)
                    rem[s][n] = __builtin_amdgcn_mfma_f64_4x4x4f64(x[s], br[n], rem[s][n], 0, 0, 0);
        }
    }
#pragma unroll
    for (int s = 0; s < ST; s++)
#pragma unroll
        for (int c = 0; c < CT; c++)
#pragma unroll
            for (int q = 0; q < 4; q++)
                T[(base + 16 * s + 4 * q + g) * KP + c * 16 + r] = acc[s][c][q];
    if constexpr (NR > 0) {
        // D_m[i][j] of this lane: row 4m + i = 4 (r / 4) + g, column r % 4
        const int row = 4 * (r >> 2) + g;
#pragma unroll
        for (int s = 0; s < ST; s++)
#pragma unroll
            for (int n = 0; n < NR; n++)
                T[(base + 16 * s + row) * KP + CT * 16 + 4 * n + (r & 3)] = rem[s][n];
        constexpr int PADC = KP - CT * 16 - 4 * NR;  // zero columns of the last tile
        if constexpr (PADC > 0)
            for (int e = lane; e < 16 * ST * PADC; e += 64)
                T[(base + e / PADC) * KP + CT * 16 + 4 * NR + e % PADC] = 0.0;
    }
}

// Whole 16-column tiles and 4-column groups for k columns padded to kp:
// R = k mod 16 <= 12 remaining columns take ceil(R / 4) groups, more a tile.
static bool f64_tiles(int k, int kp, int &ct, int &nr)
{
    ct = k / 16;
    const int R = k % 16;
    nr = (R > 0 && R <= 12) ? (R + 3) / 4 : 0;
    if (R > 12) ct++;
    return k >= 1 && 16 * (ct + (nr > 0 ? 1 : 0)) == kp && kp <= 64;
}

#define DPCA_F64_SHAPES(X)                                                                  \
    X(1, 0) X(2, 0) X(3, 0) X(4, 0) X(0, 1) X(0, 2) X(0, 3) X(1, 1) X(1, 2) X(1, 3) X(2, 1) \
        X(2, 2) X(2, 3) X(3, 1) X(3, 2) X(3, 3)

// SNP tiles per wave of k_za_f64: 4 with two or more whole column tiles
// where that form keeps >= 2 workgroups per CU (the runtime's occupancy,
// cached per shape), else 2 (also without a device)
static int za_f64_tiles(int ct, int nr)
{
    static std::atomic<int> cache[5][4];  // 0 unknown, else the tile count
    if (ct < 2 || ct > 4 || nr < 0 || nr > 3) return 2;
    const int have = cache[ct][nr].load(std::memory_order_relaxed);
    if (have) return have;
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
#define X(C_, N_)                                                                           \
    if (ct == C_ && nr == N_) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_za_f64w<C_, N_, 4>, 256, 0);
    DPCA_F64_SHAPES(X)
#undef X
    if (e != hipSuccess) return 2;
    const int t = n >= 2 ? 4 : 2;
    cache[ct][nr].store(t, std::memory_order_relaxed);
    return t;
}

int za_f64_rows_per_wg(int k, int kp)
{
    int ct, nr;
    return f64_tiles(k, kp, ct, nr) ? 64 * za_f64_tiles(ct, nr) : 128;
}

hipError_t launch_za_f64(hipStream_t st, const uint8_t *packed, int64_t rows, int64_t ldp,
                         const double *zval, const double *Q, int k, int kp, double *T,
                         int64_t seg_bytes, int nseg, double *Tseg)
{
    int ct, nr;
    if (!f64_tiles(k, kp, ct, nr)) return hipErrorInvalidValue;
    const int stl = za_f64_tiles(ct, nr);
    if (rows % (64 * stl) != 0 || seg_bytes % 64 != 0 || nseg < 1 || (nseg > 1 && !Tseg) ||
        (int64_t)nseg * seg_bytes < ldp)
        return hipErrorInvalidValue;
    dim3 grid((unsigned)(rows / (64 * stl)), (unsigned)nseg);
    double *out = nseg > 1 ? Tseg : T;
#define X(C_, N_)                                                                           \
    if (ct == C_ && nr == N_) {                                                              \
        if (stl == 4)                                                                        \
            k_za_f64w<C_, N_, 4><<<grid, 256, 0, st>>>(packed, ldp, zval, Q, out, seg_bytes, rows); \
        else                                                                                 \
            k_za_f64<C_, N_><<<grid, 256, 0, st>>>(packed, ldp, zval, Q, out, seg_bytes, rows);    \
    }
    DPCA_F64_SHAPES(X)
#undef X
    hipError_t e = hipGetLastError();
    // segments summed in order (deterministic)
    if (e == hipSuccess && nseg > 1) e = launch_reduce_parts(st, Tseg, nseg, rows * kp, T, 0);
    return e;
}

// ------------------------------------------------------- FP64  Z^T T part --
// D(16 samples x 16 rhs) += A(16 samples x 4 SNPs) B(4 SNPs x 16 rhs).
// A workgroup owns 256 samples (64 bytes of every row) and a contiguous SNP
// range (split sp); each wave owns 64 samples = 4 tiles.  64-SNP chunks of
// packed bytes, value tables and T rows are staged in LDS.  Columns as in
// k_za_f64: CT whole tiles, NR 4-column groups on v_mfma_f64_4x4x4f64 (block
// m = sample group r / 4), zero padding columns.
template <int CT, int NR>
__global__ __launch_bounds__(256) void k_zt_f64(const uint8_t *__restrict__ packed,
                                                int64_t rows, int64_t rows_per_split,
                                                int64_t ldp, const double *__restrict__ zval,
                                                const double *__restrict__ T,
                                                double *__restrict__ part, int64_t Mpad)
{
    constexpr int KP = (CT + (NR > 0 ? 1 : 0)) * 16;
    __shared__ uint8_t pk[64][64];
    __shared__ double zl[64][4];
    __shared__ double tl[64][KP];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int r = lane & 15, g = lane >> 4;
    const int64_t mb = (int64_t)blockIdx.x * 64;  // byte offset of this sample tile
    const int sp = blockIdx.y;
    const int64_t sbeg = sp * rows_per_split;
    int64_t send = sbeg + rows_per_split;
    if (send > rows) send = rows;
    d4 acc[4][CT > 0 ? CT : 1];
    double rem[4][NR > 0 ? NR : 1];
#pragma unroll
    for (int a = 0; a < 4; a++) {
#pragma unroll
        for (int c = 0; c < CT; c++) acc[a][c] = d4{0, 0, 0, 0};
#pragma unroll
        for (int n = 0; n < NR; n++) rem[a][n] = 0.0;
    }

    for (int64_t s0 = sbeg; s0 < send; s0 += 64) {
        __syncthreads();
        {
            const int rr = tid >> 2, pp = tid & 3;
            *reinterpret_cast<uint4 *>(&pk[rr][pp * 16]) =
                *reinterpret_cast<const uint4 *>(packed + (s0 + rr) * ldp + mb + pp * 16);
            zl[rr][pp] = zval[4 * (s0 + rr) + pp];
        }
        for (int e = tid; e < 64 * KP; e += 256) tl[e / KP][e % KP] = T[s0 * KP + e];
        __syncthreads();
#pragma unroll 4
        for (int step = 0; step < 16; step++) {
            const int sl = 4 * step + g;
            double b[CT > 0 ? CT : 1], br[NR > 0 ? NR : 1];
#pragma unroll
            for (int c = 0; c < CT; c++) b[c] = tl[sl][c * 16 + r];
#pragma unroll
            for (int n = 0; n < NR; n++) br[n] = tl[sl][CT * 16 + 4 * n + (r & 3)];
#pragma unroll
            for (int a = 0; a < 4; a++) {
                const int i = wave * 64 + a * 16 + r;
                const uint32_t code = (pk[sl][i >> 2] >> (2 * (i & 3))) & 3u;
                const double x = zl[sl][code];
#pragma unroll
                for (int c = 0; c < CT; c++)
                    acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, b[c], acc[a][c], 0, 0, 0);
#pragma unroll
                for (int n = 0; n < NR; n++)
                    rem[a][n] = __builtin_amdgcn_mfma_f64_4x4x4f64(x, br[n], rem[a][n], 0, 0, 0);
            }
        }
    }
    double *out = part + ((int64_t)sp * Mpad + 4 * mb + wave * 64) * KP;
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < CT; c++)
#pragma unroll
            for (int q = 0; q < 4; q++) out[(a * 16 + 4 * q + g) * KP + c * 16 + r] = acc[a][c][q];
    if constexpr (NR > 0) {
        // D_m[i][j] of this lane: sample 4m + i = 4 (r / 4) + g of the tile,
        // column r % 4 of the group
        const int row = 4 * (r >> 2) + g;
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int n = 0; n < NR; n++)
                out[(a * 16 + row) * KP + CT * 16 + 4 * n + (r & 3)] = rem[a][n];
        constexpr int PADC = KP - CT * 16 - 4 * NR;
        if constexpr (PADC > 0)
            for (int e = lane; e < 64 * PADC; e += 64)
                out[(e / PADC) * KP + CT * 16 + 4 * NR + e % PADC] = 0.0;
    }
}

int zt_f64_blocks_per_cu(int k, int kp)
{
    int ct, nr, n = 0;
    if (!f64_tiles(k, kp, ct, nr)) return 0;
    hipError_t e = hipErrorInvalidValue;
#define X(C_, N_)                                                                           \
    if (ct == C_ && nr == N_) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_zt_f64<C_, N_>, 256, 0);
    DPCA_F64_SHAPES(X)
#undef X
    return e == hipSuccess ? n : 0;
}

hipError_t launch_zt_f64(hipStream_t st, const uint8_t *packed, int64_t rows, int64_t ldp,
                         const double *zval, const double *T, int k, int kp, int splits,
                         double *part)
{
    int ct, nr;
    if (rows % 64 != 0 || ldp % 64 != 0 || splits < 1 || !f64_tiles(k, kp, ct, nr))
        return hipErrorInvalidValue;
    int64_t rps = round_up((rows + splits - 1) / splits, 64);
    dim3 grid((unsigned)(ldp / 64), (unsigned)splits);
    int64_t Mpad = 4 * ldp;
#define X(C_, N_)                                                                           \
    if (ct == C_ && nr == N_)                                                                \
        k_zt_f64<C_, N_><<<grid, 256, 0, st>>>(packed, rows, rps, ldp, zval, T, part, Mpad);
    DPCA_F64_SHAPES(X)
#undef X
    return hipGetLastError();
}

// out[e] (+)= sum over splits, in split order (deterministic).
hipError_t launch_reduce_parts(hipStream_t st, const double *part, int splits, int64_t n,
                               double *out, int accumulate)
{
    return launch_sum_parts_f64(st, part, splits, n, out, accumulate);
}

}  // namespace dpca
