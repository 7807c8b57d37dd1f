// Synthetic genotype generator: host preparation of the per-SNP population
// frequencies (Balding-Nichols: p ~ U(pmin, pmax), p_k ~ Beta(p(1-F)/F,
// (1-p)(1-F)/F)) and Dirichlet(1,...,1) admixture weights, plus the device
// kernel and host loop that turn them into packed PLINK codes.
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <thread>
#include <vector>

#include "dpca_gen.h"
#include "dpca_internal.h"
#include "k_common.h"

namespace dpca {

namespace {

// Counter-based stream for the host-side draws (one stream per SNP / sample).
struct Stream {
    uint64_t s;
    uint64_t next() { return gen_mix(s += 0x9E3779B97F4A7C15ull); }
    double uniform()  // (0, 1)
    {
        return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    }
    double normal()  // Box-Muller
    {
        double u1 = uniform(), u2 = uniform();
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
    // Marsaglia & Tsang (2000); a < 1 via Gamma(a+1) * U^(1/a).
    double gamma(double a)
    {
        if (a < 1.0) {
            double g = gamma(a + 1.0);
            return g * pow(uniform(), 1.0 / a);
        }
        const double d = a - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
        for (;;) {
            double x, v;
            do {
                x = normal();
                v = 1.0 + c * x;
            } while (v <= 0.0);
            v = v * v * v;
            const double u = uniform();
            if (u < 1.0 - 0.0331 * x * x * x * x) return d * v;
            if (log(u) < 0.5 * x * x + d * (1.0 - v + log(v))) return d * v;
        }
    }
};

uint32_t prob_to_thr(double p)
{
    if (!(p > 0.0)) return 0u;
    if (p >= 1.0) return 0xFFFFFFFFu;
    double t = floor(p * 4294967296.0);
    return t >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)t;
}

template <class F>
void parallel_for(int64_t n, int threads, F f)
{
    if (threads < 1) threads = 1;
    if (n < 4096 || threads == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    int64_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; t++) {
        int64_t a = t * chunk, b = a + chunk < n ? a + chunk : n;
        if (a >= b) break;
        th.emplace_back([=] { f(a, b); });
    }
    for (auto &x : th) x.join();
}

int host_threads()
{
    unsigned h = std::thread::hardware_concurrency();
    if (h == 0) h = 4;
    return h > 16 ? 16 : (int)h;
}

}  // namespace

int gen_prepare(int M, int64_t N, int npop, double fst, double admix_frac, double miss,
                uint64_t seed, double pmin, double pmax, GenSpec *spec, GenTables *tab)
{
    if (M <= 0 || N <= 0 || npop < 1 || fst < 0 || fst >= 1 || admix_frac < 0 ||
        admix_frac >= 1 || miss < 0 || miss >= 1 || !(pmin > 0) || !(pmax < 1) || pmin > pmax)
        return -1;
    spec->M = M;
    spec->N = N;
    spec->npop = npop;
    int n_admix = (int)llround(admix_frac * M);
    spec->m_pure = M - n_admix;
    if (spec->m_pure < npop) return -1;
    spec->miss_thr = prob_to_thr(miss);
    spec->seed = seed;
    tab->thr = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)N * npop);
    tab->wts = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)(n_admix > 0 ? n_admix : 1) * npop);
    if (!tab->thr || !tab->wts) return -1;
    parallel_for(N, host_threads(), [&](int64_t a, int64_t b) {
        for (int64_t s = a; s < b; s++) {
            Stream r{gen_mix(seed ^ (0xA0761D6478BD642Full * (uint64_t)(s + 1)))};
            const double p = pmin + (pmax - pmin) * r.uniform();
            for (int k = 0; k < npop; k++) {
                double pk = p;
                if (fst > 0) {
                    const double sh = (1.0 - fst) / fst;
                    const double x = r.gamma(p * sh), y = r.gamma((1.0 - p) * sh);
                    pk = (x + y > 0) ? x / (x + y) : p;
                }
                tab->thr[s * npop + k] = prob_to_thr(pk);
            }
        }
    });
    for (int j = 0; j < n_admix; j++) {
        Stream r{gen_mix(seed ^ 0xE7037ED1A0B428DBull ^ (0x8EBC6AF09C88C6E3ull * (uint64_t)(j + 1)))};
        std::vector<double> e(npop);
        double tot = 0;
        for (int k = 0; k < npop; k++) {
            e[k] = -log(r.uniform());  // Exp(1) -> Dirichlet(1,...,1)
            tot += e[k];
        }
        int used = 0, kmax = 0;
        for (int k = 0; k < npop; k++) {
            int w = (int)floor(e[k] / tot * 65536.0);
            tab->wts[(int64_t)j * npop + k] = (uint16_t)(w > 65535 ? 65535 : w);
            used += tab->wts[(int64_t)j * npop + k];
            if (e[k] > e[kmax]) kmax = k;
        }
        tab->wts[(int64_t)j * npop + kmax] += (uint16_t)(65536 - used);
    }
    return 0;
}

void gen_free(GenTables *tab)
{
    free(tab->thr);
    free(tab->wts);
    tab->thr = nullptr;
    tab->wts = nullptr;
}

void gen_rows_host(const GenSpec &g, const GenTables &tab, int64_t snp0, int64_t rows,
                   uint8_t *out, int threads)
{
    const int64_t np = (g.M + 3) / 4;
    parallel_for(rows, threads, [&](int64_t a, int64_t b) {
        for (int64_t r = a; r < b; r++) {
            const int64_t s = snp0 + r;
            const uint32_t *thr = tab.thr + s * g.npop;
            uint8_t *o = out + r * np;
            for (int64_t byte = 0; byte < np; byte++) {
                uint32_t v = 0;
                for (int j = 0; j < 4; j++) {
                    const int i = (int)(4 * byte + j);
                    if (i < g.M) v |= gen_code(g, s, i, gen_threshold(g, thr, tab.wts, i)) << (2 * j);
                }
                o[byte] = (uint8_t)v;  // padding codes of the last byte stay 00
            }
        }
    });
}

// One thread per packed byte of the padded device layout (grid-stride);
// bytes past the real samples hold missing codes.
__global__ void k_generate(GenSpec g, const uint32_t *__restrict__ thr,
                           const uint16_t *__restrict__ wts, int64_t snp0, int64_t rows,
                           uint8_t *__restrict__ dst, int64_t ldp)
{
    DPCA_EW_LOOP(idx, rows * ldp) {
        const int64_t r = idx / ldp, byte = idx - r * ldp;
        const int64_t s = snp0 + r;
        uint32_t v = 0;
        for (int j = 0; j < 4; j++) {
            const int i = (int)(4 * byte + j);
            const uint32_t c = (i < g.M && s < g.N)
                                   ? gen_code(g, s, i, gen_threshold(g, thr + s * g.npop, wts, i))
                                   : 1u;
            v |= c << (2 * j);
        }
        dst[idx] = (uint8_t)v;
    }
}

hipError_t launch_generate(hipStream_t st, const GenSpec &g, const uint32_t *thr,
                           const uint16_t *wts, int64_t snp0, int64_t rows, uint8_t *dst,
                           int64_t ldp)
{
    const int64_t n = rows * ldp;
    if (n == 0) return hipSuccess;
    k_generate<<<ew_blocks(n), 256, 0, st>>>(g, thr, wts, snp0, rows, dst, ldp);
    return hipGetLastError();
}

}  // namespace dpca
