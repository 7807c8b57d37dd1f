/*
 * ORACLE — test infrastructure, not product code.
 *
 * CPU restatement of the bit-exact pieces of DistPCA's randomized subspace
 * iteration path: the Omega generator, the PLINK 2-bit decode, and the per-SNP
 * standardisation with its cached 4-entry table.  Only tests/, the smoke check
 * and bench.py's cpu_baseline leg load this library, and only as the checker.
 * The linear algebra of the restatement (dgemm / dgeqrf / dorgqr / dsyev) is in
 * oracle.py, which calls the host LAPACK through numpy/scipy.
 *
 * Every function names the reference file:line it restates.  How each one is
 * pinned (known answers, compiled reference sources, closed forms) is listed in
 * the header of oracle.py.
 *
 * Build: oracle/Makefile  (gcc -O2, no fast-math, no FMA contraction).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define ORC_PI 3.14159265 /* src/gennorm.h:23 — deliberately not M_PI */

/* ------------------------------------------------------------------ RNG -- */

/* Park-Miller minimal standard generator, Schrage's factorisation
 * (src/gennorm.c:44-72).  A positive seed sets the state and returns 0; any
 * other seed advances the state one step and returns state/m. The state is
 * passed explicitly here instead of the reference's function-static. */
typedef struct { long x; } orc_rng;

double orc_rand_val(orc_rng *st, int seed)
{
    const long a = 16807, m = 2147483647, q = 127773, r = 2836;
    if (seed > 0) {
        st->x = seed;
        return 0.0;
    }
    long hi = st->x / q, lo = st->x % q;
    long t = a * lo - r * hi;
    st->x = (t > 0) ? t : t + m;
    return (double)st->x / m;
}

/* One-sided Box-Muller (src/gennorm.c:21-42): retry while the uniform is 0,
 * r = sqrt(-2 ln u), theta = 2*PI*u2, value = r*cos(theta)*sd + mean. */
double orc_norm2(orc_rng *st, double mean, double sd)
{
    double u = 0.0;
    while (u == 0.0) u = orc_rand_val(st, 0);
    double r = sqrt(-2.0 * log(u));
    double theta = 0.0;
    while (theta == 0.0) theta = 2.0 * ORC_PI * orc_rand_val(st, 0);
    double x = r * cos(theta);
    return (x * sd) + mean;
}

/* Omega, M x k row-major (src/distpca.cpp:384-397).  The reference seeds each
 * element with the C int expression jj*M + kk; that product overflows int for
 * M >= 46341 and wraps modulo 2^32 on the reference's toolchain, after which
 * non-positive seeds just continue the chain.  The wrap is computed here in
 * unsigned arithmetic and reinterpreted as int32, which is exactly what the
 * compiled reference does (checked against oracle/_ref). */
void orc_omega(int M, int k, double *out)
{
    orc_rng st = {0};
    for (int jj = 0; jj < M; jj++) {
        for (int kk = 0; kk < k; kk++) {
            uint32_t s = (uint32_t)jj * (uint32_t)M + (uint32_t)kk;
            orc_rand_val(&st, (int)(int32_t)s);
            out[(size_t)jj * k + kk] = orc_norm2(&st, 0.0, 1.0);
        }
    }
}

/* Park-Miller known answer: seed 1, n steps (n = 10000 gives 1043618065). */
long orc_lcg_after(int seed, int n)
{
    orc_rng st = {0};
    orc_rand_val(&st, seed);
    for (int i = 0; i < n; i++) orc_rand_val(&st, 0);
    return st.x;
}

/* -------------------------------------------------------------- decode -- */

/* First-touch decode (src/utilities.cpp:351-414): 2-bit codes, lowest bits =
 * lowest sample index, mapped through {0,3,1,2} (3 = missing).  Emits 4*n
 * values, padding codes included. */
void orc_decode(uint8_t *out, const uint8_t *in, unsigned n)
{
    static const uint8_t map[4] = {0, 3, 1, 2};
    for (unsigned i = 0; i < n; i++)
        for (int j = 0; j < 4; j++)
            out[4 * i + j] = map[(in[i] >> (2 * j)) & 3];
}

/* Cached-pass decode (src/utilities.cpp:417-480): codes mapped through
 * {3,1,2,0}, an index into the per-SNP 4-entry table. */
void orc_decode_precomp(uint8_t *out, const uint8_t *in, unsigned n)
{
    static const uint8_t map[4] = {3, 1, 2, 0};
    for (unsigned i = 0; i < n; i++)
        for (int j = 0; j < 4; j++)
            out[4 * i + j] = map[(in[i] >> (2 * j)) & 3];
}

/* --------------------------------------------------------- standardise -- */

/* src/utilities.cpp:272-306.  sum and count over non-missing entries, avg =
 * sum/ctr, P = avg/2, sd = sqrt(2P(1-P)) (binomial sd), z = (x-avg)/sd*s with
 * missing -> 0 and NaN -> 0.  The sum is of small integers, so its value does
 * not depend on the reference's OpenMP reduction order. */
void orc_standardize(double *normX, const uint8_t *g, int n, double *avg_out,
                     double *sd_out, double inv_sqrtN)
{
    double sum = 0;
    unsigned int ctr = 0;
    for (int i = 0; i < n; i++) {
        if (g[i] != 3) {
            sum += (double)g[i];
            ctr++;
        }
    }
    double avg = sum / ctr;
    double P = avg / 2.0;
    double sd = sqrt(2.0 * P * (1 - P));
    *avg_out = avg;
    *sd_out = sd;
    for (int i = 0; i < n; i++) {
        double x = (double)g[i];
        double z = (x == 3) ? 0 : (x - avg) / sd * inv_sqrtN;
        if (isnan(z)) z = 0;
        normX[i] = z;
    }
}

/* Decode + standardise one block of SNP rows into a dense FP64 block
 * (src/methods.cpp:814-845 and 876-905).  First touch of a SNP decodes with
 * {0,3,1,2}, standardises, and caches (0-avg)/sd, (1-avg)/sd, (2-avg)/sd and
 * 0 in the {3,2,0,1} slots of norm_precomp when sd > 1e-9; later passes decode
 * with {3,1,2,0} and multiply the cached entry by inv_sqrtN.
 *   packed : rows * np bytes (SNP-major), np = ceil(M/4)
 *   out    : rows * M doubles, row-major
 *   scratch: 4*np bytes + M doubles */
void orc_decode_block(const uint8_t *packed, int rows, int M, uint64_t np,
                      int64_t snp0, double inv_sqrtN, double *norm_precomp,
                      uint8_t *seen, double *out, uint8_t *dec, double *tmp)
{
    for (int row = 0; row < rows; row++) {
        const uint8_t *snp = packed + (uint64_t)row * np;
        int64_t gi = snp0 + row;
        double *o = out + (size_t)row * M;
        if (!seen[gi]) {
            double avg, sd;
            orc_decode(dec, snp, (unsigned)np);
            orc_standardize(tmp, dec, M, &avg, &sd, inv_sqrtN);
            if (sd > 1e-9) {
                norm_precomp[4 * gi + 3] = (0 - avg) / sd;
                norm_precomp[4 * gi + 2] = (1 - avg) / sd;
                norm_precomp[4 * gi + 0] = (2 - avg) / sd;
                norm_precomp[4 * gi + 1] = 0;
            }
            seen[gi] = 1;
            memcpy(o, tmp, sizeof(double) * (size_t)M);
        } else {
            orc_decode_precomp(dec, snp, (unsigned)np);
            for (int i = 0; i < M; i++)
                o[i] = norm_precomp[4 * gi + dec[i]] * inv_sqrtN;
        }
    }
}

/* Per-SNP statistics of a whole SNP-major matrix, in the reference's own
 * arithmetic (src/utilities.cpp:288-291 and src/methods.cpp:824-830):
 * integer sum and non-missing count, avg, sd, and the standardised value for
 * each 2-bit code, zval[4*s + code] = ((g - avg)/sd) * inv_sqrtN (code 01,
 * missing, and sd <= 1e-9 give 0). */
void orc_snp_stats(const uint8_t *packed, int64_t N, int M, int64_t *sum,
                   int64_t *ctr, double *avg, double *sd, double *zval)
{
    uint64_t np = (uint64_t)(M + 3) / 4;
    double inv_sqrtN = 1.0 / sqrt((double)N);
    for (int64_t s = 0; s < N; s++) {
        const uint8_t *row = packed + (uint64_t)s * np;
        int64_t sm = 0, c = 0;
        for (int i = 0; i < M; i++) {
            int code = (row[i >> 2] >> (2 * (i & 3))) & 3;
            if (code == 1) continue;
            c++;
            sm += (code == 0) ? 0 : (code == 2) ? 1 : 2;
        }
        double a = (double)sm / (double)(unsigned int)c;
        double P = a / 2.0;
        double d = sqrt(2.0 * P * (1 - P));
        sum[s] = sm;
        ctr[s] = c;
        avg[s] = a;
        sd[s] = d;
        double *z = zval + 4 * s;
        if (d > 1e-9) {
            z[0] = (0 - a) / d * inv_sqrtN;
            z[1] = 0;
            z[2] = (1 - a) / d * inv_sqrtN;
            z[3] = (2 - a) / d * inv_sqrtN;
        } else {
            z[0] = z[1] = z[2] = z[3] = 0;
        }
    }
}
