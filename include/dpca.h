/*
 * libdpca — C ABI of the MI355X-native randomized block subspace-iteration
 * PCA over PLINK .bed genotype matrices.
 *
 * The reference (CEID-HPCLAB/DistPCA) has no library interface: its boundary
 * is the DistPCA.exe process (src/distpca.cpp:47-644) and, inside it, the C++
 * call BlockSubspaceIter_MPI_OOC_double_buffering(const char*, double*,
 * logistics*) (src/methods.h:49, src/methods.cpp:702-1135).  Each entry point
 * below names the reference code it replaces.  Plain pointers and sizes only;
 * every call returns 0 on success or a negative code, and dpca_last_error()
 * holds the message (the CLI prints it and exits 1, as the reference does).
 * Nothing in the library aborts the process.
 *
 * Matrices are row-major, M x k (samples x right-hand sides), exactly as the
 * reference's RHS2/ARHS buffers.  One context per host thread.
 */
#ifndef DPCA_H
#define DPCA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dpca_ctx dpca_ctx;

enum { DPCA_FP64 = 0, DPCA_MIXED = 1 };

/* Run parameters.  Field meaning and defaults follow struct logistics and
 * initialize_structure (src/structures.h:54-102, src/utilities.cpp:195-232):
 * nsv 10, nrhs 20, power 1, crit 1, tol 1e-3, miter 100, bsize 0 (auto),
 * verbose 1, fwrite 0, fullsvd 0, memory 0, prefix "noprefix". */
typedef struct dpca_params {
    int nsv, nrhs, power, crit;
    double tol;
    int bsize, miter, verbose, fwrite, fullsvd;
    double memory;          /* -memory, GB (0 = RAM/10 as the reference)    */
    char prefix[1024];      /* -prefix                                       */
    int precision;          /* DPCA_FP64 (plain form) or DPCA_MIXED          */
    int device;             /* HIP device ordinal                            */
    long long device_cap;   /* bytes of HBM the packed matrix may use (0 =
                               no cap); above it the matrix streams from
                               pinned host memory in bsize-SNP blocks        */
    int strict_passes;      /* 1 (default): power+1 data passes per outer
                               iteration exactly as the reference; 0: from
                               the second iteration on, the first pass is
                               (A A^T Q) X of the previous iteration (same
                               mathematics, one pass fewer per iteration;
                               out-of-core variant only)                     */
    int host_mode;          /* where a streamed matrix lives between passes:
                               DPCA_HOST_PINNED (default; falls back to
                               DPCA_HOST_PAGEABLE when the pin fails),
                               DPCA_HOST_PAGEABLE (pageable memory staged
                               through two pinned block buffers) or
                               DPCA_HOST_FILE (dpca_open only: every pass
                               reads the .bed file again, as the reference's
                               MPI_File_iread_at loop, src/methods.cpp:795-811) */
    int slices;             /* mixed form: int8 slices of the FP64 operand of
                               each product, 2 or 3 (0 = the default, 3);
                               2 issues a third fewer MFMAs at ~100x the
                               pass error of 3                               */
    int zt_image;           /* mixed form, HBM-resident matrix: 1 (default)
                               keeps a second copy of the packed matrix in
                               the tile order of the Z^T T kernel, which then
                               reaches LDS by DMA (results bitwise the same
                               as without it); used where device_cap (if set)
                               holds both copies and the device has the room.
                               0: never                                      */
} dpca_params;

enum { DPCA_HOST_PINNED = 0, DPCA_HOST_PAGEABLE = 1, DPCA_HOST_FILE = 2 };

/* Timers and counters of one run.  The t_* fields are the reference's
 * TIME_2_* stage timers (src/structures.h:87-88), in seconds; `wall` is the
 * wall clock of dpca_run itself. */
typedef struct dpca_report {
    int M;
    long long N;
    int nsv, nrhs, power, bsize, crit;
    int total_its;          /* blockPower_total_its                          */
    int passes;             /* data passes (fetches) performed               */
    int incore;             /* 1 when bsize >= N (SubspaceIteration_MPI)     */
    int streamed;           /* 1 when the matrix streamed from host memory   */
    int qr_shifted;         /* CholQR needed its shifted fallback            */
    double t_generate_rhs, t_load, t_mm, t_mm_a, t_mm_at, t_gs, t_rr, t_other;
    double wall;
    double last_crit;       /* last criterion value (crit 0/2) or count (1)  */
    double t_h2d;           /* streamed: seconds of host-to-device copies    */
    double h2d_bytes;       /* streamed: bytes copied host-to-device         */
    int host_mode;          /* streamed: the DPCA_HOST_* mode in use         */
    int qr_repaired;        /* rank-deficient bases completed (k >= rank)     */
    double resident_bytes;  /* streamed: leading blocks kept in HBM after the
                               first pass (within device_cap)                 */
    int graph_its;          /* outer iterations replayed from captured HIP
                               graphs (resident contexts; the stage timers
                               of those are split in the proportions of the
                               eager first iteration)                        */
    int zt_image;           /* 1 when Z^T T read the tile-order image        */
} dpca_report;

/* Defaults of initialize_structure (src/utilities.cpp:195-232). */
int dpca_params_init(dpca_params *p);

/* Replaces the file handling of main (src/distpca.cpp:96-189, 239-352):
 * reads <bfile>.fam / .bim (newline-terminated lines count, as
 * src/distpca.cpp:137-147), opens <bfile>.bed, applies the bsize and
 * NSV/NRHS clamps (src/distpca.cpp:239-352) and stages the packed matrix in
 * HBM (or in pinned host memory when it exceeds device_cap). */
int dpca_open(const char *bfile, const dpca_params *p, dpca_ctx **out);

/* Same as dpca_open for a packed matrix already in host memory in .bed row
 * layout (N rows of ceil(M/4) bytes, no magic).  IDs are "<i>" for samples. */
int dpca_open_packed(const uint8_t *packed, int M, long long N, const dpca_params *p,
                     dpca_ctx **out);

/* Synthetic Balding-Nichols genotypes generated directly in HBM
 * (no reference counterpart; used for the documented workloads). */
typedef struct dpca_gen_spec {
    int M;
    long long N;
    int npop;
    double fst, admix_frac, missing;
    unsigned long long seed;
    double pmin, pmax;      /* ancestral allele frequency range              */
} dpca_gen_spec;
int dpca_open_synthetic(const dpca_gen_spec *g, const dpca_params *p, dpca_ctx **out);
/* Writes <prefix>.bed/.bim/.fam with the same generator on the host. */
int dpca_write_synthetic(const char *prefix, const dpca_gen_spec *g);

/* The bsize and NSV/NRHS resolution of src/distpca.cpp:239-352 for one rank
 * (auto block size from -memory, or from the host's MemTotal / 10 when
 * memory is 0; then the clamps), without opening anything or touching a
 * device: *out is *in with bsize, nsv, nrhs and power resolved as dpca_open
 * resolves them.  Same error returns as dpca_open's parameter checks. */
int dpca_resolve(int M, long long N, const dpca_params *in, dpca_params *out);

/* Resolved sizes after the clamps. */
int dpca_get_dims(dpca_ctx *c, int *M, long long *N, int *nsv, int *nrhs, int *bsize);

/* Reference-exact Omega, M x k row-major: the loop of src/distpca.cpp:384-397
 * over rand_val / norm2 (src/gennorm.c:21-72), including the int32 seed
 * wrap for M >= 46341.  Host-side (its log/cos are the host libm's). */
int dpca_omega(int M, int k, double *out);

/* Per-SNP statistics of the resident matrix, i.e. what standardize and the
 * norm_precomp cache compute (src/utilities.cpp:272-306,
 * src/methods.cpp:819-843): integer sum and non-missing count, and
 * zval[4*s + code] = ((g - avg)/sd) * (1/sqrt(N)) for the 2-bit code
 * (00 -> g=0, 10 -> g=1, 11 -> g=2, 01 missing -> 0).  Any pointer may be
 * NULL. */
int dpca_snp_stats(dpca_ctx *c, long long *sum, long long *ctr, double *zval);

/* Decoded codes of rows [snp0, snp0+rows) through the {0,3,1,2} table of
 * decode_plink_sse2 (src/utilities.cpp:351-414): out is rows x M bytes. */
int dpca_decode(dpca_ctx *c, long long snp0, long long rows, uint8_t *out);

/* Packed rows back in .bed layout (rows x ceil(M/4) bytes). */
int dpca_copy_packed(dpca_ctx *c, long long snp0, long long rows, uint8_t *out);

/* One data pass AAtQ = A A^T Q = Z^T (Z Q): the lambda _compute_matvec_buff
 * (src/methods.cpp:786-936).  Q, AAtQ: M x nrhs host buffers. */
int dpca_pass(dpca_ctx *c, const double *Q, double *AAtQ);

/* The subspace iteration, BlockSubspaceIter_MPI_OOC_double_buffering
 * (src/methods.cpp:938-1075) or, when bsize >= N, SubspaceIteration_MPI
 * (src/methods.cpp:24-267).  rhs_inout: M x nrhs, Omega on entry (NULL:
 * generated with dpca_omega), the Ritz vectors on return.  sigma: nsv
 * singular values (sqrt of the Ritz values), descending.  report may be
 * NULL. */
int dpca_run(dpca_ctx *c, double *rhs_inout, double *sigma, dpca_report *report);

/* Per-iteration criterion values of the last run (crit 0/2: value, crit 1:
 * converged count), as printed at -verbose 2 (src/methods.cpp:1025-1053).
 * Returns the number written (<= cap). */
int dpca_crit_history(dpca_ctx *c, double *vals, int cap);

/* <prefix>_singularValues.txt and <prefix>_leftSingularVectors.txt in the
 * formats of src/methods.cpp:1077-1118. */
int dpca_write_outputs(dpca_ctx *c, const char *prefix, const double *rhs,
                       const double *sigma);

/* Orthonormalisation unit (replaces LAPACKE_dgeqrf + dorgqr,
 * src/methods.cpp:947-948): CholQR2 on the device, Q = A R^{-1}.  A, Q:
 * m x k host buffers; returns 1 in *shifted when the shifted fallback ran. */
int dpca_orthonormalize(const double *A, long long m, int k, double *Q, int *shifted,
                        int device);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): dpca_orthonormalize on buffers already on the device,
 * with its intermediates (Gram partial sums, Gram matrix, R^{-1}, fallback
 * flags) in a static 8.4 MB device array per GPU instead of allocations.
 * A, Q: m x kp row-major device buffers, kp = k rounded up to 16, columns
 * k..kp-1 zero; A is overwritten (it holds the second CholQR pass), Q gets
 * the orthonormal factor.  Limits: k <= 64, m x kp < 2^31 (else -1 and a
 * message).  stream: a hipStream_t (NULL: the null stream).  With shifted
 * NULL the call only enqueues work (no synchronisation, so it may be
 * captured); otherwise it synchronises the stream, stores whether the shifted
 * fallback ran and returns -4 if even that failed.  Calls that overlap on one
 * GPU share the array and race.  Allocates nothing: no hipMalloc, hipFree or
 * allocating library call. */
int dpca_orthonormalize_noalloc(double *A, long long m, int k, double *Q, int *shifted,
                                int device, void *stream);

/* Symmetric eigensolver unit (replaces LAPACKE_dsyev('V','U') plus the column
 * reversal, src/methods.cpp:980-1000): upper triangle of the k x k B;
 * eigenvalues descending in w, unit eigenvectors as the columns of X. */
int dpca_syev(const double *B, int k, double *w, double *X, int device);

/* Measurement hook (no reference counterpart): runs npass passes of the
 * configured form on Q (M x nrhs) and returns event-timed averages on the
 * library's stream, in milliseconds: ms[0] whole pass, ms[1] the Z Q kernel,
 * ms[2] the Z^T T kernel, ms[3] everything else in the pass (operand
 * splits, partial reductions), ms[4] host wall clock per pass (launch to
 * stream drain), ms[5] host-to-device copy time per pass (streamed
 * matrices; it overlaps the kernels), ms[6] kernel time of the small pass
 * kernels, ms[7] the DPCA_HOST_* mode of a streamed matrix (-1 when
 * resident), ms[8] bytes copied host-to-device per pass (not a time).
 * ms must hold 9 values. */
int dpca_bench_pass(dpca_ctx *c, const double *Q, int npass, double *ms);

void dpca_close(dpca_ctx *c);
const char *dpca_last_error(void);

/* Device bytes a context with these sizes and parameters would allocate
 * (no reference counterpart; used for HBM admission). */
int dpca_footprint(int M, long long N, const dpca_params *p, long long *bytes);

/* -fullSVD diagnostic (src/distpca.cpp:449-635, src/utilities.cpp:234-270):
 * the leading nsv singular values and left singular vectors of the whole
 * standardised matrix (from the eigendecomposition of its Gram matrix on the
 * device, rocBLAS + rocSOLVER loaded at first use), compared with a finished
 * run's sigma (nsv) and Ritz vectors rhs (M x nrhs, first nsv columns):
 * true_sigma[nsv], true_U[M x nsv], sigma_relerr[j] = |true - sigma| / true,
 * vec_relerr[i x nsv + j] = ||U_ij| - |Uhat_ij|| / (|U_ij| + 1e-12),
 * cos_values[j] = (Uhat^T U)_jj.  Needs M x N x 8 bytes of device memory. */
typedef struct dpca_fullsvd_stats {
    double seconds;          /* TIME_2_TRUE_SVD                               */
    double frob_norm_angle;  /* ||Uhat^T U - I||_F with |.| on the diagonal   */
    double cos_error;        /* ||Uhat^T U - I||_2, same diagonal convention  */
    int info;                /* eigensolver status, 0 = success               */
} dpca_fullsvd_stats;
int dpca_full_svd(dpca_ctx *c, const double *rhs, const double *sigma, double *true_sigma,
                  double *true_U, double *sigma_relerr, double *vec_relerr, double *cos_values,
                  dpca_fullsvd_stats *st);
/* <prefix>_realLeftSingularVectors.txt, _realSingularValues.txt,
 * _singvecs_accuracy.txt, _cosineValues.txt in the reference's formats
 * (src/distpca.cpp:573-620). */
int dpca_write_fullsvd(const char *prefix, int M, int nsv, const double *true_U,
                       const double *true_sigma, const double *vec_relerr,
                       const double *cos_values);

/* ---- Projection and serving (no reference counterpart: DistPCA stops at
 * the left singular vectors, src/methods.cpp:1072-1118) ------------------ */

typedef struct dpca_model dpca_model;

/* SNP loadings W = Z U diag(1/sigma^2) (N x nsv, FP64) from a finished run
 * on context c (rhs = the run's M x nrhs Ritz vectors, sigma its nsv singular
 * values), with the training per-SNP table.  The model stays resident in HBM
 * and computes in the context's form (FP64 or mixed).  Projecting a training
 * sample returns its row of U. */
int dpca_model_from_run(dpca_ctx *c, const double *rhs, const double *sigma, int nsv,
                        dpca_model **out);
int dpca_model_info(dpca_model *m, long long *N, int *nsv, long long *bytes);
/* Scores (n x nsv) of n new samples given SNP-major like a .bed body: N rows
 * of ceil(n/4) bytes, same SNP order as the training matrix; missing calls
 * contribute 0 (mean imputation with the training means). */
int dpca_model_project(dpca_model *m, const uint8_t *packed, int n, double *scores);
void dpca_model_free(dpca_model *m);

typedef struct dpca_server dpca_server;
typedef struct dpca_server_params {
    int device;
    long long hbm_budget;   /* bytes the server may hold on the device       */
    int batch_capacity;     /* samples per projection batch                  */
    int workers;            /* concurrent cohort-PCA jobs (host threads)     */
} dpca_server_params;

enum { DPCA_TICKET_PROJECT = 1, DPCA_TICKET_JOB = 2 };

typedef struct dpca_ticket_info {
    int kind, status;       /* status 2 = done, < 0 error code               */
    long long n_out;        /* values written (n x nsv scores, or nsv sigma) */
    double t_submit, t_start, t_done;   /* seconds since server creation     */
    long long batch;        /* projection: batch id                          */
    int batch_samples;      /* projection: samples in that batch             */
    long long footprint;    /* job: admitted HBM bytes                       */
} dpca_ticket_info;

typedef struct dpca_server_counters {
    long long batches, batch_samples, projections_done, jobs_done, jobs_rejected,
        budget_free;
} dpca_server_counters;

int dpca_server_params_init(dpca_server_params *p);
/* model may be NULL (jobs only).  Requests that arrive while a batch is on
 * the device join the next batch (continuous batching); the batch's grouped
 * projection is a captured HIP graph replayed per batch.  Jobs are admitted
 * first-fit by dpca_footprint against the free budget and run concurrently
 * on their own streams. */
int dpca_server_create(const dpca_server_params *p, dpca_model *model, dpca_server **out);
/* The payload is read when its batch runs: keep it alive until the wait. */
int dpca_server_submit_project(dpca_server *s, const uint8_t *packed, int n, long long *ticket);
/* A cohort-PCA job on a synthetic cohort generated in HBM. */
int dpca_server_submit_job(dpca_server *s, const dpca_gen_spec *spec, const dpca_params *p,
                           long long *ticket);
/* Blocks until the ticket is done; copies up to cap results into out. */
int dpca_server_wait(dpca_server *s, long long ticket, double *out, long long cap,
                     dpca_ticket_info *info);
int dpca_server_stats(dpca_server *s, dpca_server_counters *st);
void dpca_server_destroy(dpca_server *s);
/* Page-locked host buffers for request payloads (full-speed H2D). */
void *dpca_host_alloc(long long bytes);
void dpca_host_free(void *p);

#ifdef __cplusplus
}
#endif
#endif /* DPCA_H */
