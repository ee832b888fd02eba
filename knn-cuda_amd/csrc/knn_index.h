// Launchers of the block-index kernels (knn_index.hip).
#pragma once

#include "knn_common.h"

namespace knnhip {

hipError_t idx_histogram(const int *assign, int n, int *counts, hipStream_t s);
hipError_t idx_prefix(const int *counts, int nb, int align, int *start, int *plain, hipStream_t s);
hipError_t idx_scatter(const int *assign, int n, int *cursor, int *perm, hipStream_t s);
hipError_t idx_centroids(const float *pm, int pm_ld, int dim, const int *perm, const int *start,
                         const int *counts, int nb, float *cent, int cent_ld, hipStream_t s);
hipError_t idx_transpose(const float *src, int ld, int nrows, int dim, float *dst, hipStream_t s);
hipError_t idx_init_centroids(const float *pm, int pm_ld, int n, int dim, int nb, float *cent,
                              int cent_ld, hipStream_t s);
hipError_t idx_gather_pack(const float *pm_src, int pm_ld, int dim, int oct, const int *perm,
                           int n_pos, int n_pad, float *packed, float *pm_dst, float *norms,
                           hipStream_t s);
hipError_t idx_radius(const float *pm, int pm_ld, int dim, const int *start, const int *counts,
                      int nb, const float *cent, int cent_ld, double *radius, hipStream_t s);
hipError_t idx_lower_bounds(const float *qpm, int pm_ld, int Q, int dim, const float *cent,
                            int cent_ld, int nb, const double *radius, float *lb, hipStream_t s);
hipError_t idx_pick_blocks(const float *score, int Q, int nb, int P, const float *lb,
                           const int *counts, int *chosen, float *lb_rest, hipStream_t s);
hipError_t idx_pair_count(const int *chosen, long long npairs, int *counts, hipStream_t s);
hipError_t idx_pair_scatter(const int *chosen, long long npairs, int P, int *cursor, int *qlist,
                            int *rows, hipStream_t s);
hipError_t idx_make_tasks(const int *qstart, const int *counts, int nb, const int *block_start,
                          const int *block_count, ScanTask *tasks, int max_tasks, hipStream_t s);
hipError_t idx_negate(const float *x, long long n, float *y, hipStream_t s);

}  // namespace knnhip
