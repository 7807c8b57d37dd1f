// Synthetic PLINK genotype generator (Balding-Nichols populations, optional
// Dirichlet admixture, missing calls).  The per-(SNP, sample) draw is pure
// integer arithmetic shared by the host file writer and the device kernel, so
// both produce the same bytes.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

namespace dpca {

struct GenSpec {
    int M;                 // samples
    int64_t N;             // SNPs
    int npop;              // populations (>= 1)
    int m_pure;            // samples [0, m_pure) belong to equal-size populations
    uint32_t miss_thr;     // P(missing) * 2^32
    uint64_t seed;
};

__host__ __device__ inline uint64_t gen_mix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Population of a non-admixed sample: equal-size contiguous blocks.
__host__ __device__ inline int gen_pop(const GenSpec &g, int i)
{
    return (int)(((int64_t)i * g.npop) / g.m_pure);
}

// Allele-1 threshold (probability * 2^32) of sample i at SNP s, from the
// per-SNP population thresholds thr[s*npop + k] and, for admixed samples, the
// 16-bit ancestry weights wts[(i - m_pure)*npop + k] (they sum to 65536).
__host__ __device__ inline uint32_t gen_threshold(const GenSpec &g, const uint32_t *thr_row,
                                                  const uint16_t *wts, int i)
{
    if (i < g.m_pure) return thr_row[gen_pop(g, i)];
    const uint16_t *w = wts + (int64_t)(i - g.m_pure) * g.npop;
    uint64_t acc = 0;
    for (int k = 0; k < g.npop; k++) acc += (uint64_t)w[k] * thr_row[k];
    uint64_t t = acc >> 16;
    return t > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)t;
}

// 2-bit PLINK code of (SNP s, sample i): 00 / 10 / 11 for 0 / 1 / 2 copies
// of the counted allele, 01 for a missing call.
__host__ __device__ inline uint32_t gen_code(const GenSpec &g, int64_t s, int i, uint32_t t)
{
    uint64_t x = gen_mix(g.seed ^ gen_mix((uint64_t)s * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull));
    x = gen_mix(x ^ ((uint64_t)(uint32_t)i * 0xD1B54A32D192ED03ull));
    uint64_t y = gen_mix(x);
    if ((uint32_t)(y >> 32) < g.miss_thr) return 1u;
    const uint32_t u1 = (uint32_t)(x >> 32), u2 = (uint32_t)x;
    const int gcount = (u1 < t) + (u2 < t);
    return gcount == 0 ? 0u : (gcount == 1 ? 2u : 3u);
}

// Host-side preparation: per-SNP thresholds and admixture weights.
struct GenTables {
    uint32_t *thr = nullptr;  // N * npop
    uint16_t *wts = nullptr;  // (M - m_pure) * npop
};

// Fills spec fields and allocates/fills the tables (host memory, malloc).
int gen_prepare(int M, int64_t N, int npop, double fst, double admix_frac, double miss,
                uint64_t seed, double pmin, double pmax, GenSpec *spec, GenTables *tab);
void gen_free(GenTables *tab);
// Write rows [snp0, snp0+rows) in .bed layout (np bytes per row) on the host.
void gen_rows_host(const GenSpec &g, const GenTables &tab, int64_t snp0, int64_t rows,
                   uint8_t *out, int threads);

}  // namespace dpca
