/*
 * ORACLE — test infrastructure.  Thin driver compiled together with the
 * reference's own src/gennorm.c and src/io.c (the only reference sources that
 * build here without MPI/MKL headers) into oracle/_ref/libref.so.  It restates
 * the Omega loop of src/distpca.cpp:389-394 around the reference's rand_val /
 * norm2, so tests can compare the oracle's generator with the reference's.
 */
double norm2(double mean, double std_dev);
double rand_val(int seed);

void ref_omega(int M, int k, double *out)
{
    for (int jj = 0; jj < M; jj++)
        for (int kk = 0; kk < k; kk++) {
            rand_val(jj * M + kk); /* int arithmetic, as in the reference */
            out[(long)jj * k + kk] = norm2(0.0, 1.0);
        }
}

double ref_rand_val(int seed) { return rand_val(seed); }
double ref_norm2(double mean, double sd) { return norm2(mean, sd); }
