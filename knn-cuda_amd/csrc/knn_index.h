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
hipError_t idx_pair_count(const int *chosen, int Q, int P, int p0, int p1, int *counts,
                          hipStream_t s);
hipError_t idx_pair_scatter(const int *chosen, int Q, int P, int p0, int p1, int *cursor,
                            int *qlist, int *rows, int row_base, hipStream_t s);
hipError_t idx_gather_columns(const float *src, long long ld, int dim, const int *list, int n,
                              float *dst, hipStream_t s);
hipError_t idx_scatter_results(const float *dsub, const int *isub, int k, int n, const int *list,
                               int Q, float *dist, int *idx, hipStream_t s);
hipError_t idx_make_tasks(const int *qstart, const int *counts, int nb, const int *block_start,
                          const int *block_count, ScanTask *tasks, int max_tasks, hipStream_t s);
hipError_t idx_negate(const float *x, long long n, float *y, hipStream_t s);
hipError_t idx_block_count(const int *fblock, const int *fcount, int nf, int *bcount,
                           hipStream_t s);
hipError_t idx_fine_start(const int *fblock, const int *fcount, int nf, int *bcursor, int *fstart,
                          hipStream_t s);
hipError_t idx_coarse_centroids(const float *fcent, int ld, int dim, const int *fine_by_block,
                                const int *fb_start, const int *fcount, int nb, float *cent,
                                hipStream_t s);
hipError_t idx_fine_lb(const float *qpm, int pm_ld, int Q, int dim, const float *qnorm,
                       const float *fcent, int nf, const float *fnorm, const double *frad,
                       float *lbf, hipStream_t s);
hipError_t idx_row_norms(const float *pm, int ld, int n, int dim, float *out, hipStream_t s);
hipError_t idx_block_lb(const float *lbf, int Q, int nf, const int *fine_by_block,
                        const int *fb_start, int nb, float *lbb, float *score, hipStream_t s);

}  // namespace knnhip
