// Internal declarations shared by the host driver and the HIP kernels of
// libdpca.  Nothing here is part of the C ABI (include/dpca.h is).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#define DPCA_CHECK(expr)                                                          \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) {                                                   \
            dpca::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #expr,    \
                            hipGetErrorString(e_));                               \
            return -2;                                                            \
        }                                                                         \
    } while (0)

struct dpca_ctx;

namespace dpca {

void set_error(const char *fmt, ...);

// Run every outer iteration of this context eagerly (no graph replay): the
// serving loop's jobs, where a replayed iteration enters the hardware queue
// as one block and delays the projections queued behind it.
void ctx_eager_iterations(::dpca_ctx *c);
// Contexts this thread opens (and footprints it estimates) from now on are
// serving jobs: the int8 products keep two tiles per wave and the minimum
// split count, so fewer of their waves are resident and the projections
// running beside them find room sooner (measured at 20 requests/s:
// projection p99 34-36 ms against 40-49 ms, job latency unchanged).
void set_job_contexts(bool on);

// ---------------------------------------------------------------- layout --
// Packed genotypes live in HBM SNP-major like the .bed file, but each row is
// padded to `ldp` bytes (a multiple of 256, i.e. 1024 samples) and the row
// count to a multiple of 1024.  Padding codes are 01 (missing), so they decode to 0
// in both numerical forms and no kernel needs a bounds check.
constexpr int kRowAlign = 1024;   // SNP rows padded to this multiple
constexpr int kLdpAlign = 256;    // bytes per padded row: multiple of this
constexpr uint8_t kPadByte = 0x55;
constexpr int kMaxKp = 64;         // right-hand sides, padded (k <= 64)

inline int64_t round_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// -------------------------------------------------------------- kernels --
// Every launcher enqueues on `st` and returns hipGetLastError().

// raw .bed rows (np bytes each) -> padded rows (ldp bytes), tail codes of
// the last real byte set to 01.
hipError_t launch_repack(hipStream_t st, const uint8_t *raw, int64_t rows, int64_t np,
                         uint8_t *dst, int64_t ldp, int M);
hipError_t launch_fill_u8(hipStream_t st, uint8_t *p, int64_t n, uint8_t v);
hipError_t launch_fill_f64(hipStream_t st, double *p, int64_t n, double v);
// M x k dense <-> Mpad x kp zero-padded right-hand sides
hipError_t launch_pad_rhs(hipStream_t st, const double *src, int M, int k, int64_t Mpad, int kp,
                          double *dst);
hipError_t launch_compact_rhs(hipStream_t st, const double *src, int M, int k, int kp,
                              double *dst);

// Per-SNP integer statistics and the FP64 value table zval[4*s + code].
// Also the mixed-form per-SNP mean mu[s] and scale dsc[s] = 1/(sd*sqrt(N))
// (0 where the reference's sd > 1e-9 guard fails).
hipError_t launch_snp_stats(hipStream_t st, const uint8_t *packed, int64_t rows,
                            int64_t ldp, double inv_sqrtN, int64_t *sum, int64_t *ctr,
                            double *zval, double *mu, double *dsc);

// Decode rows to reference codes {0,1,2,3=missing} (first-touch table).
// Dense FP64 standardised rows (rows x M, SNP-major) for the full SVD.
hipError_t launch_decode_f64(hipStream_t st, const uint8_t *packed, int64_t rows, int64_t ldp,
                             int M, const double *zval, double *out);
// Top nsv eigenpairs of the Gram matrix of B (M x N, column-major, ld M,
// on the device) through rocBLAS / rocSOLVER loaded at first use: sigma
// (descending) and the left singular vectors U (M x nsv, row-major) of B,
// both on the host; *info is the eigensolver's status.
int dense_top_svd(hipStream_t st, const double *B, int M, int64_t N, int nsv, double *sigma,
                  double *U, int *info);
hipError_t launch_decode_codes(hipStream_t st, const uint8_t *packed, int64_t rows,
                               int64_t ldp, int M, uint8_t *codes);

// FP64 form, T = Z Q over all padded rows.  T: rows x kp.  With nseg > 1 the
// samples are split into nseg segments of seg_bytes packed bytes (more
// workgroups for short blocks), partials in Tseg (nseg x rows x kp), summed
// in segment order into T.
// k columns are computed (kp = k rounded up to 16; the rest written as 0).
hipError_t launch_za_f64(hipStream_t st, const uint8_t *packed, int64_t rows, int64_t ldp,
                         const double *zval, const double *Q, int k, int kp, double *T,
                         int64_t seg_bytes, int nseg, double *Tseg);
// SNP rows one k_za_f64 workgroup covers (128 or 256)
int za_f64_rows_per_wg(int k, int kp);
// FP64 form, partial[sp] = Z^T T over split sp of the rows (k columns as above).
hipError_t launch_zt_f64(hipStream_t st, const uint8_t *packed, int64_t rows, int64_t ldp,
                         const double *zval, const double *T, int k, int kp, int splits,
                         double *part);
// workgroups of k_zt_f64 for k columns resident per CU (0 if unknown)
int zt_f64_blocks_per_cu(int k, int kp);
// out[e] (+)= sum over p < nparts of part[p * n + e] in FP64, in an order
// fixed by (nparts, n, e) alone.
hipError_t launch_sum_parts_f64(hipStream_t st, const double *part, int nparts, int64_t n,
                                double *out, int accumulate);
hipError_t launch_sum_parts_f32(hipStream_t st, const float *part, int nparts, int64_t n,
                                double *out, int accumulate);
// out (+)= sum_sp part[sp]; n = Mpad*kp
hipError_t launch_reduce_parts(hipStream_t st, const double *part, int splits, int64_t n,
                               double *out, int accumulate);

// Mixed form: exact int8 codes x two or three int8 slices of the FP64 operand,
// int32 accumulation; see k_mixed.hip.
// T = sum of the nseg FP32 partials of T (rows a multiple of 128); with
// smax non-null also the per-split maxima that launch_uw8 needs (splits of
// rps rows), so that it can be called with have_max.
hipError_t launch_tprep(hipStream_t st, const float *Tpart, int nseg, int64_t rows, int kp,
                        double *T, const double *mu, const double *dsc, int64_t rps,
                        unsigned *smax);
hipError_t launch_reduce_f32(hipStream_t st, const float *part, int splits, int64_t n,
                             double *out, int accumulate);
// Column tiles of the int8 forms: the NS slices (2 or 3) of k right-hand
// columns as 16-wide MFMA tiles.  F whole column tiles take NS tiles each;
// the R columns left over share RT = ceil(NS R / 16) tiles when that is
// fewer than NS, else they are padded to a whole tile.  kp() is the padded
// column count of the FP64 buffers (k rounded up to 16 either way).
struct ColTiles {
    int NS, F, R, RT;
    int vt() const { return NS * F + RT; }
    int kp() const { return 16 * (F + (RT > 0 ? 1 : 0)); }
};
inline ColTiles col_tiles(int k, int ns)
{
    ColTiles t{ns, k / 16, k % 16, 0};
    t.RT = (ns * t.R + 15) / 16;
    if (t.RT >= ns) {
        t.F += 1;
        t.R = 0;
        t.RT = 0;
    }
    return t;
}
// Z^T T on int8 MFMA: U = d T, W = -d mu T as ns int8 slices per (SNP split
// of rps rows, column) with a power-of-two scale from smax.
// UW8: rows/8 x vt x 16 words of 16 bytes; smax: ceil(rows/rps) x kp
// (computed here unless have_max: then already computed by launch_tprep).
hipError_t launch_uw8(hipStream_t st, const double *T, int64_t rows, int k, int ns, int kp,
                      const double *mu, const double *dsc, int64_t rps, unsigned *smax, void *UW8,
                      bool have_max);
// img: the tile-order image of the same rows (launch_zt_image, 8-wave
// workgroups), or null to stage the codes from packed.
hipError_t launch_zt_i8(hipStream_t st, const uint8_t *packed, const uint8_t *img, int64_t rows,
                        int64_t ldp, const void *UW8, const unsigned *smax, int k, int ns,
                        int sa_tiles, int waves, int64_t rows_per_split, int nsplit, float *part);
// Tile-order image of rows x ldp packed bytes for k_zt_i8's DMA form (rows x
// ldp bytes at img; rows a multiple of 128, 8-wave workgroups).
hipError_t launch_zt_image(hipStream_t st, const uint8_t *packed, int64_t rows, int64_t ldp,
                           int sa_tiles, int waves, uint8_t *img);
// Z Q on int8 MFMA: codes as exact h / m rows, Q as ns int8 slices per
// (sample segment of 4*seg_bytes samples, column) with a power-of-two scale
// from qmax.  Q8: Mpad/16 x vt x 16 words of 16 bytes; qmax: nseg x kp.
hipError_t launch_q8(hipStream_t st, const double *Q, int64_t Mpad, int k, int ns, int kp,
                     int64_t seg_bytes, unsigned *qmax, void *Q8);
hipError_t launch_za_i8(hipStream_t st, const uint8_t *packed, int64_t rows, int64_t ldp,
                        const void *Q8, const unsigned *qmax, const double *mu,
                        const double *dsc, int k, int ns, int st_tiles, int waves,
                        int64_t seg_bytes, int nseg, float *Tpart);
// Both int8 products run workgroups of `waves` = 8 waves (two per SIMD) or 4
// (one per SIMD, twice the registers: larger tiles per wave).  Tiles per
// wave (sample tiles of k_zt_i8 for a batch of ldb bytes, SNP tiles of
// k_za_i8): the largest form whose accumulators fit the registers, unless a
// smaller one doubles the workgroups resident per CU (needs the current
// device; without one the largest form).  *_valid says whether a tile count
// is instantiated for these columns (the tuning knobs).
int zt_i8_tiles(int k, int ns, int64_t ldb, int waves);
int za_i8_tiles(int k, int ns, int waves);
bool zt_i8_valid(int k, int ns, int sa_tiles, int64_t ldb, int waves);
bool za_i8_valid(int k, int ns, int st_tiles, int waves);
// workgroups of k_zt_i8 with sa_tiles resident per CU (0 if unknown)
int zt_i8_blocks_per_cu(int k, int ns, int sa_tiles, int waves);
// packed bytes per row that one k_zt_i8 workgroup covers
int64_t zt_i8_tile_bytes(int sa_tiles, int waves);

// Tall-skinny helpers, FP64.  X, Y: m x kp row-major.
hipError_t launch_gram(hipStream_t st, const double *X, const double *Y, int64_t m, int kp,
                       double *part, int nparts, double *G);
hipError_t launch_chol_inv(hipStream_t st, const double *G, int k, int kp, double shift_mult,
                           int mode, double *Rinv, int *flags, double *cinfo);
hipError_t launch_replace_cols(hipStream_t st, double *A, int M, int64_t Mpad, int k, int kp,
                               const double *cinfo);
hipError_t launch_apply(hipStream_t st, const double *X, const double *W, int64_t m, int kp,
                        double *Y);
hipError_t launch_syev(hipStream_t st, const double *B, int k, int kp, double *w, double *V,
                       int *status);
hipError_t launch_col_dots(hipStream_t st, const double *X, const double *Y, int64_t m,
                           int kp, int ncols, double *part, int nparts, double *dots);
hipError_t launch_colsum(hipStream_t st, const double *X, int64_t m, int kp, double *part,
                         int nparts, double *out);

// Projection batches and loadings.
hipError_t launch_scatter_cols(hipStream_t st, const uint8_t *src, int64_t rows, int64_t nb,
                               int n_valid, uint8_t *dst, int64_t ldb, int64_t off);
hipError_t launch_fill_cols(hipStream_t st, uint8_t *dst, int64_t rows, int64_t ldb, int64_t c0,
                            int64_t c1);
hipError_t launch_loadings(hipStream_t st, const double *T, int64_t rows, int kp, int nsv,
                           const double *inv_s2, double *W);

// Per-iteration state of the convergence test (device resident).
struct IterState {
    double sigma[64];
    double sigma_old[64];
    double delta, delta_prev, crit;
    int conv, stop;
};
hipError_t launch_crit(hipStream_t st, const double *w, int nsv, int crit, double tol, int ii,
                       const double *dots, IterState *s);

// Synthetic generator (device side).  See dpca_gen.h.
struct GenSpec;
hipError_t launch_generate(hipStream_t st, const GenSpec &g, const uint32_t *thr,
                           const uint16_t *wts, int64_t snp0, int64_t rows, uint8_t *dst,
                           int64_t ldp);

}  // namespace dpca
