// Internal: a resident projection model (SNP loadings + training per-SNP
// statistics) shared by the model entry points and the serving loop.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dpca_internal.h"

struct dpca_model {
    int dev = 0;
    int precision = 0;
    int ns = 3;                // mixed form: int8 slices of the loadings
    int64_t N = 0, Npad = 0;   // SNPs of the training matrix (rows of a payload)
    int nsv = 0, k = 0, kp = 0;  // k: columns of the training context (its column tiles)
    double *zval = nullptr;    // Npad x 4, training table (FP64 form)
    double *mu = nullptr, *dsc = nullptr;
    double *W = nullptr;       // Npad x kp loadings Z U diag(1/sigma^2)
    uint4 *UW8 = nullptr;      // mixed form: int8 slices of d W and -d mu W
    unsigned *smax = nullptr;  // mixed form: their per (split, column) maxima
    size_t bytes = 0;          // device footprint of the above
};

namespace dpca {

// Projection of samples held in a packed batch matrix (Npad rows x ldb
// bytes, 4 samples per byte, padding codes 01): scores[(4*ldb) x kp] in FP64,
// scores[i][j] = sum_s z_si W_sj with z standardised by the training table.
// `part` holds the split partials (FP64 form: splits x 4*ldb x kp doubles;
// mixed: nsplit x 4*ldb x kp floats; the int8 slices of the mixed form are
// scaled per split, so the split is also the model's slice-scale group).  Launch parameters are fixed per
// (model, ldb) so the launches can be captured in a graph; the SNP split is
// fixed per model (batch-invariant reduction order).
constexpr int64_t kProjSplitRowsF64 = 4096;  // SNPs per split, FP64 form
constexpr int64_t kProjSplitRowsMx = 2048;   // SNPs per split, mixed form
inline int64_t proj_rps_mx(int64_t Npad) { return Npad < kProjSplitRowsMx ? Npad : kProjSplitRowsMx; }
struct ProjPlan {
    int64_t ldb = 0;
    int splits = 1;            // FP64 form
    int sa_tiles = 2;          // mixed form
    int64_t rps = 0;
    int nsplit = 1;
    size_t part_bytes = 0;
};
ProjPlan plan_projection(const dpca_model *m, int64_t ldb);
hipError_t launch_projection(hipStream_t st, const dpca_model *m, const ProjPlan &pl,
                             const uint8_t *batch, void *part, double *scores);

}  // namespace dpca
