// Dense FP64 algebra on the device for the -fullSVD diagnostic
// (src/distpca.cpp:449-635): the Gram matrix of the whole standardised
// matrix through rocBLAS and its eigendecomposition through rocSOLVER.
// libdpca does not link either library; they are loaded at the first call,
// so every other entry point works without them.
//
// The reference computes an economy SVD of the M x N matrix (dgesvd).  Here
// the left singular vectors and singular values come from the symmetric
// eigendecomposition of the smaller Gram matrix (M x M, or N x N with
// U = B V diag(1/sigma)).  For the leading pairs the diagnostic compares,
// the eigenvalues carry an absolute error of about eps * sigma_1^2, i.e. a
// relative error of eps * (sigma_1 / sigma_j)^2 in sigma_j^2.
#include <dlfcn.h>
#include <math.h>
#include <stdio.h>

#include <mutex>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "dpca_internal.h"

namespace dpca {

namespace {

struct DenseLib {
    decltype(&rocblas_create_handle) create = nullptr;
    decltype(&rocblas_destroy_handle) destroy = nullptr;
    decltype(&rocblas_set_stream) set_stream = nullptr;
    decltype(&rocblas_dgemm) dgemm = nullptr;
    decltype(&rocsolver_dsyevd) syevd = nullptr;
    bool ok = false;
    char why[256] = "";
};

void *open_any(const char *const *names)
{
    for (int i = 0; names[i]; i++)
        if (void *h = dlopen(names[i], RTLD_NOW | RTLD_GLOBAL)) return h;
    return nullptr;
}

const DenseLib &dense_lib()
{
    static DenseLib lib;
    static std::once_flag once;
    std::call_once(once, [] {
        const char *blas_names[] = {"librocblas.so.5", "librocblas.so",
                                    "/opt/rocm/lib/librocblas.so", nullptr};
        const char *solver_names[] = {"librocsolver.so.0", "librocsolver.so",
                                      "/opt/rocm/lib/librocsolver.so", nullptr};
        void *b = open_any(blas_names);
        void *s = b ? open_any(solver_names) : nullptr;
        if (!b || !s) {
            snprintf(lib.why, sizeof lib.why, "cannot load %s: %s", b ? "rocSOLVER" : "rocBLAS",
                     dlerror());
            return;
        }
        lib.create = (decltype(lib.create))dlsym(b, "rocblas_create_handle");
        lib.destroy = (decltype(lib.destroy))dlsym(b, "rocblas_destroy_handle");
        lib.set_stream = (decltype(lib.set_stream))dlsym(b, "rocblas_set_stream");
        lib.dgemm = (decltype(lib.dgemm))dlsym(b, "rocblas_dgemm");
        lib.syevd = (decltype(lib.syevd))dlsym(s, "rocsolver_dsyevd");
        lib.ok = lib.create && lib.destroy && lib.set_stream && lib.dgemm && lib.syevd;
        if (!lib.ok) snprintf(lib.why, sizeof lib.why, "rocBLAS / rocSOLVER symbols missing");
    });
    return lib;
}

}  // namespace

int dense_top_svd(hipStream_t st, const double *B, int M, int64_t N, int nsv, double *sigma,
                  double *U, int *info)
{
    const DenseLib &L = dense_lib();
    if (!L.ok) {
        set_error("full SVD: %s", L.why);
        return -5;
    }
    if (N > 0x7fffffff || nsv < 1 || nsv > M || nsv > N) {
        set_error("full SVD: unsupported shape %d x %lld (nsv %d)", M, (long long)N, nsv);
        return -1;
    }
    const int n = (int)N;
    const bool over_samples = M <= n;   // Gram over the smaller dimension
    const int g = over_samples ? M : n;
    double *G = nullptr, *D = nullptr, *E = nullptr, *Ud = nullptr;
    int *dinfo = nullptr;
    rocblas_handle h = nullptr;
    int rc = 0;
    auto fail = [&](const char *what, int code) {
        set_error("full SVD: %s failed (%d)", what, code);
        rc = -2;
    };
    if (hipMalloc((void **)&G, (size_t)g * g * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&D, (size_t)g * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&E, (size_t)g * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&dinfo, sizeof(int)) != hipSuccess ||
        (!over_samples && hipMalloc((void **)&Ud, (size_t)M * nsv * sizeof(double)) != hipSuccess)) {
        set_error("full SVD: device allocation of the %d x %d Gram matrix failed", g, g);
        rc = -3;
    }
    rocblas_status s = rocblas_status_success;
    const double one = 1.0, zero = 0.0;
    if (!rc && (s = L.create(&h)) != rocblas_status_success) fail("rocblas_create_handle", s);
    if (!rc && (s = L.set_stream(h, st)) != rocblas_status_success) fail("rocblas_set_stream", s);
    if (!rc) {
        // B is M x N column-major: G = B B^T (samples) or B^T B (SNPs)
        s = over_samples
                ? L.dgemm(h, rocblas_operation_none, rocblas_operation_transpose, M, M, n, &one, B,
                          M, B, M, &zero, G, M)
                : L.dgemm(h, rocblas_operation_transpose, rocblas_operation_none, n, n, M, &one, B,
                          M, B, M, &zero, G, n);
        if (s != rocblas_status_success) fail("rocblas_dgemm", s);
    }
    if (!rc && (s = L.syevd(h, rocblas_evect_original, rocblas_fill_upper, g, G, g, D, E, dinfo)) !=
                   rocblas_status_success)
        fail("rocsolver_dsyevd", s);
    if (!rc && !over_samples) {
        // U = B V_top: the eigenvectors of the nsv largest eigenvalues are the
        // last nsv columns of G (ascending order)
        s = L.dgemm(h, rocblas_operation_none, rocblas_operation_none, M, nsv, n, &one, B, M,
                    G + (size_t)(g - nsv) * g, g, &zero, Ud, M);
        if (s != rocblas_status_success) fail("rocblas_dgemm", s);
    }
    std::vector<double> w(g), vec;
    if (!rc) {
        vec.resize((size_t)(over_samples ? g : M) * nsv);
        const double *src = over_samples ? G + (size_t)(g - nsv) * g : Ud;
        if (hipMemcpyAsync(w.data(), D, (size_t)g * sizeof(double), hipMemcpyDeviceToHost, st) !=
                hipSuccess ||
            hipMemcpyAsync(vec.data(), src, vec.size() * sizeof(double), hipMemcpyDeviceToHost,
                           st) != hipSuccess ||
            hipMemcpyAsync(info, dinfo, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            set_error("full SVD: copy of the eigenpairs failed");
            rc = -2;
        }
    }
    if (!rc) {
        // column (nsv - 1 - j) of vec is the j-th largest pair
        for (int j = 0; j < nsv; j++) {
            const double lam = w[g - 1 - j];
            sigma[j] = lam > 0 ? sqrt(lam) : 0.0;
            const double *col = vec.data() + (size_t)(nsv - 1 - j) * M;
            const double inv = (!over_samples && sigma[j] > 0) ? 1.0 / sigma[j] : 1.0;
            for (int i = 0; i < M; i++) U[(size_t)i * nsv + j] = col[i] * inv;
        }
    }
    if (h) (void)L.destroy(h);
    void *bufs[] = {G, D, E, Ud, dinfo};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    return rc;
}

}  // namespace dpca
