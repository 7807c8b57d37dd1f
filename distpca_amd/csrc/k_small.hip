// Small dense FP64 kernels of the subspace iteration: tall-skinny Gram
// products, Cholesky + triangular inverse (CholQR2), the M x k by k x k
// product, the k x k symmetric eigensolver (parallel cyclic Jacobi) and the
// column dot products of the MEV criterion.  All reductions run in a fixed
// order, so results do not depend on scheduling.  k <= 64 throughout.
#include "dpca_internal.h"
#include "k_common.h"

#include <algorithm>

namespace dpca {

// ------------------------------------------------------------------- Gram --
// part[p][a*kp + b] = sum_{i in rows of p} X[i][a] * Y[i][b]; 32-row tiles of
// X and Y are staged in LDS; thread t owns entries t, t+256, ...
template <int KP>
__global__ __launch_bounds__(256) void k_gram_part(const double *__restrict__ X,
                                                   const double *__restrict__ Y, int64_t m,
                                                   int64_t rpp, double *__restrict__ part)
{
    constexpr int NE = (KP * KP + 255) / 256;
    __shared__ double xs[32][KP + 1];
    __shared__ double ys[32][KP + 1];
    const int tid = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * rpp;
    int64_t r1 = r0 + rpp;
    if (r1 > m) r1 = m;
    double acc[NE];
#pragma unroll
    for (int u = 0; u < NE; u++) acc[u] = 0;
    for (int64_t rb = r0; rb < r1; rb += 32) {
        __syncthreads();
        for (int e = tid; e < 32 * KP; e += 256) {
            const int rr = e / KP, cc = e % KP;
            const int64_t i = rb + rr;
            xs[rr][cc] = (i < r1) ? X[i * KP + cc] : 0.0;
            ys[rr][cc] = (i < r1) ? Y[i * KP + cc] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NE; u++) {
            const int e = tid + 256 * u;
            if (e < KP * KP) {
                const int a = e / KP, b = e % KP;
                double s = acc[u];
#pragma unroll 8
                for (int rr = 0; rr < 32; rr++) s += xs[rr][a] * ys[rr][b];
                acc[u] = s;
            }
        }
    }
#pragma unroll
    for (int u = 0; u < NE; u++) {
        const int e = tid + 256 * u;
        if (e < KP * KP) part[(int64_t)blockIdx.x * KP * KP + e] = acc[u];
    }
}

// out[e] (+)= sum over p < nparts of part[p * n + e], in FP64.  W waves per
// workgroup (W from nparts only): wave w sums the partials p = w, w + W, ...
// for 64 consecutive e, then the W sums are added in wave order, so the
// result depends on (nparts, n, e) alone, never on timing or on the grid.
template <typename T>
__global__ __launch_bounds__(1024) void k_sum_parts(const T *__restrict__ part, int nparts,
                                                    int64_t n, int W, double *__restrict__ out,
                                                    int accumulate)
{
    __shared__ double red[16][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t e = (int64_t)blockIdx.x * 64 + lane;
    double s = 0;
    if (e < n) {
#pragma unroll 4
        for (int p = w; p < nparts; p += W) s += (double)part[(int64_t)p * n + e];
    }
    red[w][lane] = s;
    __syncthreads();
    if (w == 0 && e < n) {
        double t = accumulate ? out[e] : 0.0;
        for (int j = 0; j < W; j++) t += red[j][lane];
        out[e] = t;
    }
}

template <typename T>
static hipError_t sum_parts(hipStream_t st, const T *part, int nparts, int64_t n, double *out,
                            int accumulate)
{
    if (n <= 0) return hipSuccess;
    // >= ~8 partials per wave, at most 16 waves
    const int W = nparts >= 128 ? 16 : std::max(1, nparts / 8);
    k_sum_parts<T><<<(unsigned)((n + 63) / 64), 64 * W, 0, st>>>(part, nparts, n, W, out,
                                                                  accumulate);
    return hipGetLastError();
}

hipError_t launch_sum_parts_f64(hipStream_t st, const double *part, int nparts, int64_t n,
                                double *out, int accumulate)
{
    return sum_parts(st, part, nparts, n, out, accumulate);
}

hipError_t launch_sum_parts_f32(hipStream_t st, const float *part, int nparts, int64_t n,
                                double *out, int accumulate)
{
    return sum_parts(st, part, nparts, n, out, accumulate);
}

// Row blocks of >= 32 rows (one LDS tile): small m uses fewer partials.
static int gram_parts(int64_t m, int nparts)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>(nparts, (m + 31) / 32));
}

hipError_t launch_gram(hipStream_t st, const double *X, const double *Y, int64_t m, int kp,
                       double *part, int nparts, double *G)
{
    nparts = gram_parts(m, nparts);
    int64_t rpp = (m + nparts - 1) / nparts;
    switch (kp) {
    case 16: k_gram_part<16><<<nparts, 256, 0, st>>>(X, Y, m, rpp, part); break;
    case 32: k_gram_part<32><<<nparts, 256, 0, st>>>(X, Y, m, rpp, part); break;
    case 48: k_gram_part<48><<<nparts, 256, 0, st>>>(X, Y, m, rpp, part); break;
    case 64: k_gram_part<64><<<nparts, 256, 0, st>>>(X, Y, m, rpp, part); break;
    default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return sum_parts(st, (const double *)part, nparts, (int64_t)kp * kp, G, 0);
}

// -------------------------------------------------- Cholesky + R^{-1} ------
// G (kp x kp, leading k x k used; its upper triangle defines the symmetric
// matrix) = R^T R with R upper triangular; writes R^{-1} (kp x kp, zero
// outside the leading block).  CholQR2 with the shifted-CholeskyQR3 fallback
// of Fukaya et al., SIAM J. Sci. Comput. 42 (2020):
//   mode 0  factor G; if a pivot is not positive, raise the diagonal by
//           shift_mult * trace(G), factor again and set flags[0] = 1;
//   mode 1  factor G;
//   mode 2  third pass: R^{-1} = I unless the shift was used or the Gram of
//           the CholQR2 result is not yet the identity.
// A failure that survives the shift sets flags[1] = 1.
// Right-looking Cholesky of the leading k x k block of A (upper triangle).
// A pivot that is not positive, or that has lost all but ~1e-14 of its
// column's diagonal (the column lies within ~1e-7 rad of the span of the
// previous ones, beyond what CholQR2 can orthonormalise), marks the factor bad
// and sets colbad[j]; that column's row of R is then zeroed beyond the
// diagonal, so the later pivots are those of the basis without it and colbad
// ends up naming a set of columns whose removal leaves the rest independent.
__device__ int chol_factor(double (*A)[kMaxKp + 1], const double *diag0, int k, int *bad,
                           int *colbad)
{
    const int tid = threadIdx.x;
    for (int j = 0; j < k; j++) {
        // every thread forms the pivot itself (two barriers per column);
        // A[j][j] is written only in the update phase, which never reads it
        const double d = A[j][j];
        const bool dj = !(d > 0.0) || !(d > 1e-14 * diag0[j]);
        const double piv = dj ? 1.0 : sqrt(d);
        for (int c = j + 1 + tid; c < k; c += blockDim.x) A[j][c] = dj ? 0.0 : A[j][c] / piv;
        __syncthreads();
        if (tid == 0) {
            if (dj) *bad = 1;
            colbad[j] = dj;
            A[j][j] = piv;
        }
        const int n = k - j - 1;
        for (int e = tid; e < n * n; e += blockDim.x) {
            const int r = j + 1 + e / n, c = j + 1 + e % n;
            if (r <= c) A[r][c] -= A[j][r] * A[j][c];
        }
        __syncthreads();
    }
    return *bad;
}

// mode 2 also runs a real third pass when the Gram of the CholQR2 result is
// not the identity to 1e-12 (flags[3] records it).  When cinfo is not NULL,
// mode 0 records the unshifted factor's dependent columns there (cinfo[j] =
// 1.0 for a dropped column j, else 0.0) and the largest diagonal of G in
// cinfo[kMaxKp], for k_replace_cols.
__global__ __launch_bounds__(256) void k_chol_inv(const double *__restrict__ G, int k,
                                                  int kp, double shift_mult, int mode,
                                                  double *__restrict__ Rinv, int *flags,
                                                  double *__restrict__ cinfo)
{
    __shared__ double A[kMaxKp][kMaxKp + 1];
    __shared__ double col[kMaxKp];
    __shared__ double diag0[kMaxKp];
    __shared__ int colbad[kMaxKp];
    __shared__ int bad, skip;
    const int tid = threadIdx.x;
    if (tid == 0) {
        bad = 0;
        skip = 0;
    }
    for (int e = tid; e < kp * kp; e += blockDim.x) {
        const int i = e / kp, j = e % kp;
        A[i][j] = (i <= j) ? G[i * kp + j] : G[j * kp + i];
    }
    __syncthreads();
    // the relative-pivot test applies to the first factorisation only
    if (tid < k) diag0[tid] = (mode == 0) ? A[tid][tid] : 0.0;
    if (mode == 2 && tid == 0 && flags[0] == 0) {
        double dev = 0;
        for (int i = 0; i < k; i++)
            for (int j = i; j < k; j++) dev = fmax(dev, fabs(A[i][j] - (i == j ? 1.0 : 0.0)));
        skip = !(dev > 1e-12);
        if (!skip) flags[3] = 1;
    }
    __syncthreads();
    if (skip) {
        for (int e = tid; e < kp * kp; e += blockDim.x)
            Rinv[e] = (e / kp == e % kp && e / kp < k) ? 1.0 : 0.0;
        return;
    }
    chol_factor(A, diag0, k, &bad, colbad);
    if (mode == 0 && cinfo) {
        if (tid < k) cinfo[tid] = colbad[tid] ? 1.0 : 0.0;
        if (tid == 0) {
            double dm = 0;
            for (int i = 0; i < k; i++) dm = fmax(dm, diag0[i]);
            cinfo[kMaxKp] = dm;
        }
    }
    if (bad && mode == 0) {
        __syncthreads();
        for (int e = tid; e < kp * kp; e += blockDim.x) {
            const int i = e / kp, j = e % kp;
            A[i][j] = (i <= j) ? G[i * kp + j] : G[j * kp + i];
        }
        __syncthreads();
        if (tid == 0) {
            double tr = 0;
            for (int i = 0; i < k; i++) tr += A[i][i];
            const double sh = shift_mult * tr;
            for (int i = 0; i < k; i++) {
                A[i][i] += sh;
                diag0[i] = 0.0;  // the shifted factor only needs positive pivots
            }
            bad = 0;
            flags[0] = 1;
        }
        __syncthreads();
        chol_factor(A, diag0, k, &bad, colbad);
    }
    // in-place inverse of the upper triangle, one column at a time (column j
    // and its diagonal are written only after every read of them)
    for (int j = 0; j < k; j++) {
        const double djj = A[j][j];
        if (tid < j) {
            double s = 0;
            for (int l = tid; l < j; l++) s += A[tid][l] * A[l][j];
            col[tid] = s;
        }
        __syncthreads();
        if (tid < j) A[tid][j] = -col[tid] / djj;
        if (tid == 0) A[j][j] = 1.0 / djj;
        __syncthreads();
    }
    for (int e = tid; e < kp * kp; e += blockDim.x) {
        const int i = e / kp, j = e % kp;
        Rinv[e] = (i < k && j < k && i <= j) ? A[i][j] : 0.0;
    }
    if (tid == 0 && bad) flags[1] = 1;
}

hipError_t launch_chol_inv(hipStream_t st, const double *G, int k, int kp, double shift_mult,
                           int mode, double *Rinv, int *flags, double *cinfo)
{
    if (kp > kMaxKp || k > kp) return hipErrorInvalidValue;
    // 256 threads: one wave alone (cheaper barriers) measured 1.4x slower
    k_chol_inv<<<1, 256, 0, st>>>(G, k, kp, shift_mult, mode, Rinv, flags, cinfo);
    return hipGetLastError();
}

// ------------------------------------------- dependent-column replacement --
// A (Mpad x kp, row-major): every column j < k that the last mode-0 factor
// dropped (cinfo[j] != 0) is overwritten by a fixed pseudo-random vector
// (splitmix64 of the row and column, uniform in [-1, 1), rows < M only; the
// padding rows stay zero) scaled to the largest column norm of A.  A basis
// whose columns span less than k dimensions (k at or above the rank of
// A A^T, e.g. k = M with centred columns) is thereby completed with
// directions in general position, as Householder QR completes it with
// directions chosen by its rounding errors.
__device__ inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

__global__ __launch_bounds__(256) void k_replace_cols(double *__restrict__ A, int M, int64_t Mpad,
                                                      int k, int kp,
                                                      const double *__restrict__ cinfo)
{
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= Mpad * kp) return;
    const int64_t i = idx / kp;
    const int j = (int)(idx % kp);
    if (j >= k || cinfo[j] == 0.0) return;
    double v = 0.0;
    if (i < M) {
        const uint64_t h = splitmix64(((uint64_t)i << 7) ^ (uint64_t)j);
        v = ((double)(h >> 11) * 0x1.0p-53 * 2.0 - 1.0) * sqrt(3.0 * cinfo[kMaxKp] / M);
    }
    A[idx] = v;
}

hipError_t launch_replace_cols(hipStream_t st, double *A, int M, int64_t Mpad, int k, int kp,
                               const double *cinfo)
{
    if (kp > kMaxKp || k > kp) return hipErrorInvalidValue;
    const unsigned nb = (unsigned)((Mpad * kp + 255) / 256);
    k_replace_cols<<<nb, 256, 0, st>>>(A, M, Mpad, k, kp, cinfo);
    return hipGetLastError();
}

// ------------------------------------------------------------ Y = X W ------
// X: m x kp, W: kp x kp (zero outside the used block), Y: m x kp.  W is
// staged in LDS; one thread per output element.
template <int KP>
__global__ __launch_bounds__(256) void k_apply(const double *__restrict__ X,
                                               const double *__restrict__ W, int64_t m,
                                               double *__restrict__ Y)
{
    __shared__ double ws[KP][KP];
    for (int e = threadIdx.x; e < KP * KP; e += 256) ws[e / KP][e % KP] = W[e];
    __syncthreads();
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= m * KP) return;
    const int64_t i = idx / KP;
    const int j = (int)(idx % KP);
    const double *x = X + i * KP;
    double s = 0;
#pragma unroll 16
    for (int l = 0; l < KP; l++) s += x[l] * ws[l][j];
    Y[idx] = s;
}

hipError_t launch_apply(hipStream_t st, const double *X, const double *W, int64_t m, int kp,
                        double *Y)
{
    const unsigned nb = (unsigned)((m * kp + 255) / 256);
    switch (kp) {
    case 16: k_apply<16><<<nb, 256, 0, st>>>(X, W, m, Y); break;
    case 32: k_apply<32><<<nb, 256, 0, st>>>(X, W, m, Y); break;
    case 48: k_apply<48><<<nb, 256, 0, st>>>(X, W, m, Y); break;
    case 64: k_apply<64><<<nb, 256, 0, st>>>(X, W, m, Y); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ------------------------------------------------- symmetric eigensolver ---
// Cyclic Jacobi with the round-robin parallel ordering (Brent & Luk): each
// round rotates n/2 disjoint pairs at once; n - 1 rounds make a sweep.  The
// rotation is Golub & Van Loan's sym.schur2.  Input: the upper triangle of B
// (kp x kp, leading k x k).  Output: eigenvalues in descending order in w[0..k)
// and the matching unit eigenvectors as the columns of V (kp x kp).
__global__ __launch_bounds__(256) void k_syev_jacobi(const double *__restrict__ B, int k,
                                                     int kp, double *__restrict__ w,
                                                     double *__restrict__ V, int *status)
{
    __shared__ double A[kMaxKp][kMaxKp + 1];
    __shared__ double U[kMaxKp][kMaxKp + 1];
    __shared__ double rc[kMaxKp / 2], rs[kMaxKp / 2];
    __shared__ int rp[kMaxKp / 2], rq[kMaxKp / 2];
    __shared__ double red[256];
    __shared__ int done;
    const int tid = threadIdx.x;
    for (int e = tid; e < k * k; e += blockDim.x) {
        const int i = e / k, j = e % k;
        A[i][j] = (i <= j) ? B[i * kp + j] : B[j * kp + i];
        U[i][j] = (i == j) ? 1.0 : 0.0;
    }
    const int n = (k % 2) ? k + 1 : k;  // one dummy index when k is odd
    const int half = n / 2;
    int sweeps = 0;
    __syncthreads();
    for (; sweeps < 40; sweeps++) {
        // convergence: off-diagonal mass against the diagonal, summed as
        // 256 strided partial sums and a fixed tree (the same additions
        // whatever the block size)
        double dia[256 / 64];  // blockDim.x >= 64
        for (int vt = tid; vt < 256; vt += blockDim.x) {
            double off = 0, dg = 0;
            for (int e = vt; e < k * k; e += 256) {
                const int i = e / k, j = e % k;
                const double v = A[i][j] * A[i][j];
                if (i == j) dg += v; else off += v;
            }
            red[vt] = off;
            dia[vt / blockDim.x] = dg;
        }
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            for (int vt = tid; vt < o; vt += blockDim.x) red[vt] += red[vt + o];
            __syncthreads();
        }
        const double offs = red[0];
        __syncthreads();
        for (int vt = tid; vt < 256; vt += blockDim.x) red[vt] = dia[vt / blockDim.x];
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            for (int vt = tid; vt < o; vt += blockDim.x) red[vt] += red[vt + o];
            __syncthreads();
        }
        if (tid == 0) done = !(offs > 1e-30 * red[0]);
        __syncthreads();
        if (done) break;
        for (int round = 0; round < n - 1; round++) {
            if (tid < half) {
                int p, q;
                if (tid == 0) {
                    p = 0;
                    q = 1 + round % (n - 1);
                } else {
                    p = 1 + (round + tid) % (n - 1);
                    q = 1 + (round + n - 1 - tid) % (n - 1);
                }
                if (p > q) { const int t = p; p = q; q = t; }
                double c = 1.0, s = 0.0;
                if (q < k) {
                    const double apq = A[p][q];
                    if (apq != 0.0) {
                        const double tau = (A[q][q] - A[p][p]) / (2.0 * apq);
                        const double t = (tau >= 0 ? 1.0 : -1.0) /
                                         (fabs(tau) + sqrt(1.0 + tau * tau));
                        c = 1.0 / sqrt(1.0 + t * t);
                        s = t * c;
                    }
                }
                rp[tid] = p;
                rq[tid] = q;
                rc[tid] = c;
                rs[tid] = s;
            }
            __syncthreads();
            // rows: A <- J^T A
            for (int e = tid; e < half * k; e += blockDim.x) {
                const int pi = e / k, j = e % k;
                const int p = rp[pi], q = rq[pi];
                if (q < k) {
                    const double c = rc[pi], s = rs[pi];
                    const double ap = A[p][j], aq = A[q][j];
                    A[p][j] = c * ap - s * aq;
                    A[q][j] = s * ap + c * aq;
                }
            }
            __syncthreads();
            // columns: A <- A J,  U <- U J
            for (int e = tid; e < half * k; e += blockDim.x) {
                const int pi = e / k, i = e % k;
                const int p = rp[pi], q = rq[pi];
                if (q < k) {
                    const double c = rc[pi], s = rs[pi];
                    const double ap = A[i][p], aq = A[i][q];
                    A[i][p] = c * ap - s * aq;
                    A[i][q] = s * ap + c * aq;
                    const double up = U[i][p], uq = U[i][q];
                    U[i][p] = c * up - s * uq;
                    U[i][q] = s * up + c * uq;
                }
            }
            __syncthreads();
        }
    }
    // descending order: rank of each diagonal entry (ties by index)
    if (tid < k) {
        const double d = A[tid][tid];
        int rank = 0;
        for (int j = 0; j < k; j++) {
            const double e = A[j][j];
            if (e > d || (e == d && j < tid)) rank++;
        }
        w[rank] = d;
        for (int i = 0; i < kp; i++) V[i * kp + rank] = (i < k) ? U[i][tid] : 0.0;
    }
    for (int e = tid; e < kp * kp; e += blockDim.x) {
        const int i = e / kp, j = e % kp;
        if (j >= k) V[e] = 0.0;
        (void)i;
    }
    if (tid == 0 && sweeps >= 40) *status = 1;
}

hipError_t launch_syev(hipStream_t st, const double *B, int k, int kp, double *w, double *V,
                       int *status)
{
    if (kp > kMaxKp || k > kp) return hipErrorInvalidValue;
    // 256 threads: one wave alone (cheaper barriers) measured 2x slower
    k_syev_jacobi<<<1, 256, 0, st>>>(B, k, kp, w, V, status);
    return hipGetLastError();
}

// ------------------------------------------------- column dot products -----
// part[p][j] = sum_{i in rows of p} X[i][j] * Y[i][j], j < ncols.
__global__ __launch_bounds__(256) void k_col_dots(const double *__restrict__ X,
                                                  const double *__restrict__ Y, int64_t m,
                                                  int kp, int ncols, int64_t rpp,
                                                  double *__restrict__ part)
{
    __shared__ double red[256];
    const int tid = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * rpp;
    int64_t r1 = r0 + rpp;
    if (r1 > m) r1 = m;
    const int rows_per_pass = 256 / kp;
    const int j = tid % kp, lr = tid / kp;
    double s = 0;
    if (lr < rows_per_pass && j < ncols)
        for (int64_t i = r0 + lr; i < r1; i += rows_per_pass)
            s += (Y ? X[i * kp + j] * Y[i * kp + j] : X[i * kp + j]);
    red[tid] = (lr < rows_per_pass) ? s : 0.0;
    __syncthreads();
    if (tid < kp) {
        double t = 0;
        for (int l = 0; l < rows_per_pass; l++) t += red[l * kp + tid];
        part[(int64_t)blockIdx.x * kp + tid] = (tid < ncols) ? t : 0.0;
    }
}

hipError_t launch_col_dots(hipStream_t st, const double *X, const double *Y, int64_t m, int kp,
                           int ncols, double *part, int nparts, double *dots)
{
    nparts = gram_parts(m, nparts);
    int64_t rpp = (m + nparts - 1) / nparts;
    k_col_dots<<<nparts, 256, 0, st>>>(X, Y, m, kp, ncols, rpp, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return sum_parts(st, (const double *)part, nparts, (int64_t)kp, dots, 0);
}

hipError_t launch_colsum(hipStream_t st, const double *X, int64_t m, int kp, double *part,
                         int nparts, double *out)
{
    return launch_col_dots(st, X, nullptr, m, kp, kp, part, nparts, out);
}

}  // namespace dpca

namespace dpca {

// ------------------------------------------------- convergence criteria ----
// sigma_j = sqrt(w_j) for the nsv largest Ritz values and the three stopping
// tests of src/methods.cpp:1012-1056 (crit 0: relative change of the sum of
// singular values; crit 1: every |(s_j - s_j_old)/s_j| <= tol; crit 2:
// 1 - (1/nsv) sum_j (q_old_j . q_new_j)^2 <= tol), evaluated on the device in
// the reference's operation order.
__global__ void k_crit(const double *__restrict__ w, int nsv, int crit, double tol, int ii,
                       const double *__restrict__ dots, IterState *s)
{
#pragma clang fp contract(off)
    if (threadIdx.x != 0) return;
    double delta = 0.0;
    for (int j = 0; j < nsv; j++) {
        s->sigma[j] = sqrt(w[j]);
        delta += s->sigma[j];
    }
    int stop = 0, conv = 0;
    double cval = 0.0;
    if (ii > 0) {
        if (crit == 0) {
            cval = fabs(s->delta_prev - delta) / delta;
            stop = cval <= tol;
        } else if (crit == 1) {
            for (int j = 0; j < nsv; j++)
                if (fabs((s->sigma[j] - s->sigma_old[j]) / s->sigma[j]) <= tol) conv++;
            cval = conv;
            stop = conv == nsv;
        } else if (crit == 2) {
            double mev = 0.0;
            for (int j = 0; j < nsv; j++) mev += dots[j] * dots[j];
            mev /= nsv;
            cval = 1.0 - mev;
            stop = cval <= tol;
        }
    }
    s->delta = delta;
    s->crit = cval;
    s->conv = conv;
    s->stop = stop;
    s->delta_prev = delta;
    if (!stop)
        for (int j = 0; j < nsv; j++) s->sigma_old[j] = s->sigma[j];
}

hipError_t launch_crit(hipStream_t st, const double *w, int nsv, int crit, double tol, int ii,
                       const double *dots, IterState *s)
{
    if (nsv > kMaxKp) return hipErrorInvalidValue;
    k_crit<<<1, 64, 0, st>>>(w, nsv, crit, tol, ii, dots, s);
    return hipGetLastError();
}

}  // namespace dpca
