// DistPCA.exe — drop-in command line for a single-rank run of the reference
// (src/distpca.cpp:47-644) on one MI355X through libdpca.
//
// Flags, defaults, clamps, console layout and output files follow the
// reference: -bfile -nsv -nrhs -power -crit -tol -bsize -miter -verbose
// -fwrite -prefix -memory -fullSVD -about.  Extra flags of this build:
// -precision fp64|mixed, -device <ordinal>, -devcap <GB of HBM for the packed
// matrix before it streams from host memory>, -strict 0 (reuse the last pass
// of an iteration as the next iteration's first: power passes per iteration
// instead of power+1), -hostmem pinned|pageable|file (where a streamed matrix
// lives between passes; file re-reads the .bed every pass), -slices 2|3 (int8
// slices of the mixed form's FP64 operand), -ztimage 0 (no tile-order copy of
// the matrix for the mixed form's Z^T T kernel).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/sysinfo.h>
#include <time.h>

#include <string>
#include <vector>

#include "dpca.h"

namespace {

enum ArgType { INT, DOUBLE, STR, NA };

// "-name value" lookup with the semantics of findarg (src/io.c:24-68):
// the first argument equal to "-name" wins; a missing value leaves the field
// unchanged; INT/DOUBLE go through atoi/atof.
int findarg(const char *name, ArgType type, void *val, int argc, char **argv)
{
    for (int i = 0; i < argc; i++) {
        if (argv[i][0] != '-') continue;
        if (strcmp(name, argv[i] + 1) != 0) continue;
        if (type == NA) return 1;
        if (i + 1 >= argc) return 0;
        switch (type) {
        case INT: *(int *)val = atoi(argv[i + 1]); return 1;
        case DOUBLE: *(double *)val = atof(argv[i + 1]); return 1;
        case STR: snprintf((char *)val, 1024, "%s", argv[i + 1]); return 1;
        default: return 0;
        }
    }
    return 0;
}

std::string stamp()
{
    time_t t = time(nullptr);
    char *s = asctime(localtime(&t));
    std::string str(s);
    if (!str.empty() && str.back() == '\n') str.pop_back();
    return "[" + str + "] ";
}

int ram_kb()
{
    struct sysinfo si;
    if (sysinfo(&si) != 0) return 0;
    return (int)((double)si.totalram * si.mem_unit / 1024.0);
}

// Console report of print_statistics (src/utilities.cpp:147-193).  The
// scripts read field 5 of the first "Total wall-clock time elapsed:" line of
// the seconds block; as in the reference it is the sum of the stage timers.
// The fetch count is the reference's total_its * (power + 1) under its own
// schedule; under -strict 0 (strict == false) it is the passes the run made.
void print_statistics(const dpca_report &r, int threads, int rkb, bool fullsvd, bool strict)
{
    time_t now = time(nullptr);
    printf("----------------------------------------------------------\n");
    printf("DistPCA software package, Version: 1.0.0\n");
    printf("Current local time and date: %s", asctime(localtime(&now)));
    printf("Total number of threads exploited: %d\n", threads);
    printf("RAM size in KBs: %d\n", rkb);
    printf("RAM size in GBs: %lf\n", rkb / 1e6);
    printf("Total number of matrix rows: %d\n", r.M);
    printf("Total number of matrix columns: %lld\n", r.N);
    printf("Total number of singular pairs sought: %d\n", r.nsv);
    printf("Total number of right-hand sides used: %d\n", r.nrhs);
    printf("Value of power: %d\n", r.power);
    printf("Total number of rows fetched from the disk per block: %d\n", r.bsize);
    if (r.bsize < r.N) {
        const int fetches = strict ? r.total_its * (r.power + 1) : r.passes;
        printf("Total number of times which the data matrix was fetched from the memory: %d\n",
               fetches);
        printf("Average amount of time elapsed per matrix fetching: %lf\n", r.t_load / fetches);
    }
    const double tot = r.t_rr + r.t_generate_rhs + r.t_load + r.t_mm + r.t_gs + r.t_other;
    const double tot_noload = tot - r.t_load;
    const double scale[2] = {1.0, 3600.0};
    const char *unit[2] = {"seconds", "hours"};
    for (int u = 0; u < 2; u++) {
        const double d = scale[u];
        printf("\n[The following times are listed in %s]\n\n", unit[u]);
        printf("Time to generate the right-hand sides matrix: %02.13f\n", r.t_generate_rhs / d);
        printf("Time to load the data matrix: %02.13f\n", r.t_load / d);
        printf("Time to perform the MM products (overall): %02.13f\n", r.t_mm / d);
        printf("Time to perform the MM products (with A): %02.13f\n", r.t_mm_a / d);
        printf("Time to perform the MM products (with A^T): %02.13f\n", r.t_mm_at / d);
        printf("Time to perform the orthonormalization: %02.13f\n", r.t_gs / d);
        printf("Time to solve the projection eigenvalue problem: %02.13f\n", r.t_rr / d);
        printf("\nTotal wall-clock time elapsed: %02.13f\n", tot / d);
        printf("Total wall-clock time elapsed (without including the %s spent on loading the "
               "matrix): %02.13f\n",
               u == 0 ? "amount of time" : "time", tot_noload / d);
    }
    // The reference prints these lines before it computes the full SVD
    // (src/distpca.cpp:446-449), so they always carry the initial zeros.
    if (r.incore && fullsvd) {
        printf("\nTotal wall-clock time to compute the full ('econ') SVD: %02.13f\n", 0.0);
        printf("||Uhat^TU - I||_F: %02.13f\n", 0.0);
        printf("||Uhat^TU - I||_2: %02.13f\n", 0.0);
    }
}

int llc_kb()
{
    // get_llc_size_kb (src/utilities.cpp:101-126)
    FILE *fp = fopen("/sys/devices/system/cpu/cpu0/cache/index3/size", "r");
    if (!fp) return -1;
    int kb = -1;
    if (fscanf(fp, "%dK", &kb) != 1) kb = -1;
    fclose(fp);
    return kb;
}

// Memory, cache and block-size lines of src/distpca.cpp:219-306 for one rank.
// p holds the flags as given; bsize is the block size the library resolved
// from them (the reference's rows_fetched).
void print_sizing(const dpca_params &p, int M, long long N, int bsize, int rkb)
{
    const double mem = p.memory == 0.0 ? rkb / 1e6 / 10.0 : p.memory;
    printf("\n%sMemory per rank: %g GB\n", stamp().c_str(), mem);
    const int llc = llc_kb();
    if (llc > 0) printf("\n%sDetected LLC size: %d KB (%d MB)\n", stamp().c_str(), llc, llc / 1024);
    if (p.bsize > 0 && p.bsize > N)
        printf("\n%sWarning: Specified block size (%d) exceeds local SNPs (%lld). Adjusting to "
               "local_N.\n",
               stamp().c_str(), p.bsize, N);
    printf("\n%sBlock size (SNPs per block): %d\n", stamp().c_str(), bsize);
    // the auxiliary estimate uses NRHS as given, before its clamps
    const double block_gb = 2.0 * bsize * (double)M * 8.0 / 1e9;
    const double aux_gb = (double)M * p.nrhs * 8.0 * 3 / 1e9;
    printf("\n%sEstimated memory per rank: %g GB (Block buffers: %g GB, Auxiliary arrays: %g GB)\n",
           stamp().c_str(), block_gb + aux_gb, block_gb, aux_gb);
    printf("\n%sBlock size (in MB): %g (typical LLC sizes: 16-64 MB per socket)\n",
           stamp().c_str(), (double)bsize * (double)M * 8.0 / (1024.0 * 1024.0));
}

// The NSV / NRHS adjustments of src/distpca.cpp:325-352, replayed on the
// flags as given to print the reference's messages; the library applies the
// same rules when it resolves the parameters.
void print_clamps(const dpca_params &p, int M, long long N)
{
    const int mn = (int)(M < N ? M : N);
    int s = p.nsv, r = p.nrhs;
    auto two_s = [&] { return 2 * s < mn ? 2 * s : mn; };
    if (s > mn) {
        s = mn;
        printf("\n%sThe value of NSV given was larger than min(M, N). Adjusting to NSV = "
               "min(M,N)...\n",
               stamp().c_str());
    }
    if (r <= 0) {
        printf("\n%sNRHS was either zero or negative. Adjusting to NRHS = min(2xNSV, "
               "min(M,N))...\n",
               stamp().c_str());
        r = two_s();
    }
    if (r > mn) {
        r = mn;
        printf("\n%sThe value of NRHS given was larger than min(M,N). Adjusting to NRHS = "
               "min(M,N)...\n",
               stamp().c_str());
    }
    if (r < s) {
        printf("\n%sNRHS can not be smaller than NSV. Adjusting to NRHS = "
               "min(2xNSV,min(M,N))...\n",
               stamp().c_str());
        r = two_s();
    }
}

}  // namespace

int main(int argc, char **argv)
{
    dpca_params p;
    dpca_params_init(&p);
    if (findarg("about", NA, nullptr, argc, argv)) {
        printf("\nDistPCA\nDistributed Out-of-Core Principal Component Analysis (PCA) for "
               "Tera-Scale Genomic Data\n(MI355X build: one GPU through libdpca)\n\n");
        return 0;
    }
    char fname[1024] = "", pname[1024] = "", prec[1024] = "";
    int pflag = findarg("prefix", STR, pname, argc, argv);
    findarg("bfile", STR, fname, argc, argv);
    findarg("nrhs", INT, &p.nrhs, argc, argv);
    findarg("nsv", INT, &p.nsv, argc, argv);
    findarg("memory", DOUBLE, &p.memory, argc, argv);
    findarg("bsize", INT, &p.bsize, argc, argv);
    findarg("verbose", INT, &p.verbose, argc, argv);
    findarg("fwrite", INT, &p.fwrite, argc, argv);
    findarg("tol", DOUBLE, &p.tol, argc, argv);
    findarg("miter", INT, &p.miter, argc, argv);
    findarg("crit", INT, &p.crit, argc, argv);
    findarg("power", INT, &p.power, argc, argv);
    findarg("fullSVD", INT, &p.fullsvd, argc, argv);
    findarg("device", INT, &p.device, argc, argv);
    findarg("slices", INT, &p.slices, argc, argv);
    int zti = 1;
    if (findarg("ztimage", INT, &zti, argc, argv)) p.zt_image = zti != 0;
    int strict = 1;
    if (findarg("strict", INT, &strict, argc, argv)) p.strict_passes = strict != 0;
    double devcap_gb = 0;
    if (findarg("devcap", DOUBLE, &devcap_gb, argc, argv)) p.device_cap = (long long)(devcap_gb * 1e9);
    char hostmem[1024] = "";
    if (findarg("hostmem", STR, hostmem, argc, argv)) {
        if (!strcmp(hostmem, "pinned"))
            p.host_mode = DPCA_HOST_PINNED;
        else if (!strcmp(hostmem, "pageable"))
            p.host_mode = DPCA_HOST_PAGEABLE;
        else if (!strcmp(hostmem, "file"))
            p.host_mode = DPCA_HOST_FILE;
        else {
            printf("\nUnknown -hostmem '%s' (pinned | pageable | file)\n", hostmem);
            return 1;
        }
    }
    if (findarg("precision", STR, prec, argc, argv)) {
        if (!strcmp(prec, "mixed"))
            p.precision = DPCA_MIXED;
        else if (!strcmp(prec, "fp64"))
            p.precision = DPCA_FP64;
        else {
            printf("\nUnknown -precision '%s' (fp64 | mixed)\n", prec);
            return 1;
        }
    }
    if (pflag) snprintf(p.prefix, sizeof p.prefix, "%s", pname);

    const std::string bfile(fname);
    printf("%sReading .fam file: %s.fam\n", stamp().c_str(), bfile.c_str());
    dpca_ctx *ctx = nullptr;
    struct timespec o0, o1;
    clock_gettime(CLOCK_MONOTONIC, &o0);
    if (dpca_open(bfile.c_str(), &p, &ctx) != 0) {
        printf("\n%s\n", dpca_last_error());
        return 1;
    }
    clock_gettime(CLOCK_MONOTONIC, &o1);
    const double t_open = (o1.tv_sec - o0.tv_sec) + 1e-9 * (o1.tv_nsec - o0.tv_nsec);
    int M, nsv, nrhs, bsize;
    long long N;
    dpca_get_dims(ctx, &M, &N, &nsv, &nrhs, &bsize);
    const int rkb = ram_kb();
    printf("\n%sNumber of individuals (M): %d\n", stamp().c_str(), M);
    printf("\n%sReading .bim file: %s.bim\n", stamp().c_str(), bfile.c_str());
    printf("\n%sNumber of markers (N): %lld\n", stamp().c_str(), N);
    printf("\n%sNumber of MPI ranks: 1\n", stamp().c_str());
    printf("\n%sSNPs per rank: ~%lld\n", stamp().c_str(), N);
    printf("\n%sDataset partitioning across MPI ranks:\n\n", stamp().c_str());
    printf("Rank 0: SNPs [0, %lld) = %lld SNPs\n", N, N);
    printf("\n%sSize of RAM (in KB): %d\n", stamp().c_str(), rkb);
    print_sizing(p, M, N, bsize, rkb);
    print_clamps(p, M, N);
    if (bsize == N) {
        // src/distpca.cpp:366-377; the library reads the .bed inside
        // dpca_open, so the time is that call's
        printf("\n%sReading .bed file (distributed across ranks)...\n", stamp().c_str());
        printf("\n%sFinished reading local data. Time: %g sec\n", stamp().c_str(), t_open);
    }
    printf("\n%sPrecision: %s on HIP device %d\n", stamp().c_str(),
           p.precision == DPCA_MIXED ? "mixed" : "fp64", p.device);
    fflush(stdout);

    std::vector<double> rhs((size_t)M * nrhs), sigma(nsv);
    dpca_report rep;
    {
        // Omega of src/distpca.cpp:384-397, timed like TIME_2_GENERATE_RHS
        struct timespec a, b;
        clock_gettime(CLOCK_MONOTONIC, &a);
        dpca_omega(M, nrhs, rhs.data());
        clock_gettime(CLOCK_MONOTONIC, &b);
        const double tg = (b.tv_sec - a.tv_sec) + 1e-9 * (b.tv_nsec - a.tv_nsec);
        if (dpca_run(ctx, rhs.data(), sigma.data(), &rep) != 0) {
            printf("\n%s\n", dpca_last_error());
            dpca_close(ctx);
            return 1;
        }
        rep.t_generate_rhs = tg;
    }
    if (p.verbose > 1) {
        printf("\nApproximate eigenvalues:\n");
        for (int j = 0; j < nsv; j++) printf("λ_%d = %.13f\n", j, sigma[j] * sigma[j]);
        printf("\nFirst 5 individuals' eigenvectors:\n\n");
        printf("%-20s", "FID");
        for (int j = 0; j < nsv; j++) printf("%-15s", ("PC" + std::to_string(j)).c_str());
        printf("\n%s\n", std::string(20 + 15 * nsv, '-').c_str());
        // the first five IIDs are printed from the output writer's table
        for (int i = 0; i < (M < 5 ? M : 5); i++) {
            printf("%-20d", i);
            for (int j = 0; j < nsv; j++) printf("%-15.6f", rhs[(size_t)i * nrhs + j]);
            printf("\n");
        }
    }
    if (p.verbose > 0) {
        const char *omp = getenv("OMP_NUM_THREADS");
        print_statistics(rep, omp ? atoi(omp) : 1, rkb, p.fullsvd == 1, p.strict_passes != 0);
    }
    if (p.fwrite == 1 && dpca_write_outputs(ctx, p.prefix, rhs.data(), sigma.data()) != 0) {
        printf("\n%s\n", dpca_last_error());
        dpca_close(ctx);
        return 1;
    }
    if (p.fullsvd == 1) {
        // accuracy diagnostic of src/distpca.cpp:449-635
        std::vector<double> ts(nsv), tu((size_t)M * nsv), se(nsv), ve((size_t)M * nsv), cv(nsv);
        dpca_fullsvd_stats fs;
        if (dpca_full_svd(ctx, rhs.data(), sigma.data(), ts.data(), tu.data(), se.data(),
                          ve.data(), cv.data(), &fs) != 0) {
            printf("\n%s\n", dpca_last_error());
            dpca_close(ctx);
            return 1;
        }
        if (p.verbose > 1) {
            printf("\nFULL SVD computed in %lf sec, info = %d\n\n", fs.seconds, fs.info);
            for (int j = 0; j < nsv; j++) printf("RelError Sigma(%d): %e\n", j, se[j]);
            for (int j = 0; j < nsv; j++) printf("Cosine(%d): %lf\n", j, cv[j]);
        }
        if (p.fwrite == 1 &&
            dpca_write_fullsvd(p.prefix, M, nsv, tu.data(), ts.data(), ve.data(), cv.data()) != 0) {
            printf("%s\n", dpca_last_error());
            dpca_close(ctx);
            return 1;
        }
    }
    dpca_close(ctx);
    return 0;
}
