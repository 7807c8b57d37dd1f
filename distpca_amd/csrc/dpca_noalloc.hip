// Allocation-free CholQR2 (dpca_orthonormalize_noalloc): the orthonormal
// factor of a tall-skinny matrix already in device memory, with every
// intermediate the kernels need (Gram partial sums, the Gram matrix, R^{-1},
// the fallback flags) in one file-scope device array instead of hipMalloc'd
// buffers, so that callers which must not allocate (code under HIP graph
// capture, serving loops) can call it.
//
// The array is per device: each GPU the library runs on holds its own copy
// (8.4 MB) until the process ends.  Every call on a device shares it, so calls
// that overlap on one device race; the array keeps the last call's contents
// and each call clears the flags it reads.
#include <math.h>

#include <algorithm>
#include <mutex>

#include "dpca.h"
#include "dpca_internal.h"

namespace dpca {

namespace {

constexpr int kParts = 256;                  // Gram partials (as the contexts use)
constexpr size_t kPartElems = (size_t)kParts * kMaxKp * kMaxKp;
constexpr size_t kMatElems = (size_t)kMaxKp * kMaxKp;
constexpr size_t kFlagElems = 8;             // 16 ints as doubles' storage
constexpr size_t kScratchElems = kPartElems + 2 * kMatElems + kFlagElems;

// Partials | G | R^{-1} | flags.  In host code this name is only a stand-in
// for the symbol: the device address comes from hipGetSymbolAddress.
static __device__ __align__(256) double scratch[kScratchElems];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};

// The device address of `scratch` on device dev (current), resolved once per
// device.
double *scratch_on(int dev)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("dpca_orthonormalize_noalloc: hipGetSymbolAddress: %s",
                      hipGetErrorString(e));
            return nullptr;
        }
        g_ptr[dev] = static_cast<double *>(p);
    }
    return g_ptr[dev];
}

}  // namespace

}  // namespace dpca

extern "C" int dpca_orthonormalize_noalloc(double *A, long long m, int k, double *Q, int *shifted,
                                           int device, void *stream)
{
    using namespace dpca;
    const int64_t kp = round_up(k, 16);
    if (!A || !Q || A == Q || k <= 0 || m < k) {
        set_error("dpca_orthonormalize_noalloc: bad arguments (A and Q distinct device "
                  "buffers, m >= k >= 1)");
        return -1;
    }
    if (k > kMaxKp || (int64_t)m * kp >= (1ll << 31)) {
        set_error("dpca_orthonormalize_noalloc: m x k = %lld x %d exceeds its limits (k <= %d, "
                  "m x roundup(k,16) < 2^31)",
                  m, k, kMaxKp);
        return -1;
    }
    if (device < 0 || device >= kMaxDevices) {
        set_error("dpca_orthonormalize_noalloc: device %d out of range", device);
        return -1;
    }
    DPCA_CHECK(hipSetDevice(device));
    double *base = scratch_on(device);
    if (!base) return -2;
    double *part = base, *G = base + kPartElems, *Rinv = G + kMatElems;
    int *flags = reinterpret_cast<int *>(Rinv + kMatElems);
    hipStream_t st = static_cast<hipStream_t>(stream);
    // same shift as the contexts (11 (m k + k (k + 1)) u) and the same three
    // passes as their non-robust path: A -> Q -> A -> Q
    const double u = ldexp(1.0, -53);
    const double shift = 11.0 * ((double)std::min<long long>(m, 0x7fffffff) * k +
                                 (double)k * (k + 1)) * u;
    // a fill kernel, not a memset node: safe inside a captured graph
    DPCA_CHECK(launch_fill_u8(st, reinterpret_cast<uint8_t *>(flags), kFlagElems * 8, 0));
    double *src = A;
    double *dst[3] = {Q, A, Q};
    for (int mode = 0; mode < 3; mode++) {
        DPCA_CHECK(launch_gram(st, src, src, m, (int)kp, part, kParts, G));
        DPCA_CHECK(launch_chol_inv(st, G, k, (int)kp, shift, mode, Rinv, flags, nullptr));
        DPCA_CHECK(launch_apply(st, src, Rinv, m, (int)kp, dst[mode]));
        src = dst[mode];
    }
    if (!shifted) return 0;  // no synchronisation: the caller may be capturing
    int fl[2] = {0, 0};
    DPCA_CHECK(hipMemcpyAsync(fl, flags, sizeof fl, hipMemcpyDeviceToHost, st));
    DPCA_CHECK(hipStreamSynchronize(st));
    *shifted = fl[0];
    if (fl[1]) {
        set_error("CholeskyQR failed even with the shifted fallback");
        return -4;
    }
    return 0;
}
