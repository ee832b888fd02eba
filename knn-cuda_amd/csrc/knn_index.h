// Launchers of the block-index kernels (knn_index.hip).
#pragma once

#include "knn_common.h"

namespace knnhip {

// Fine balls per block on average (the fine k-means has this many centres
// per block).
constexpr int kBallsPerBlock = 4;
// Workgroups of a k-means++ seeding step.
constexpr int kKppGroups = 256;
// Largest padded dimension the block index supports (LDS of the seeding
// kernel).
constexpr int kIndexMaxLd = 4096;
// Largest block count (LDS of the grouping kernel).
constexpr int kIndexMaxBlocks = 32768;

hipError_t idx_histogram(const int *assign, int n, int *counts, hipStream_t s);
hipError_t idx_prefix(const int *counts, int nb, int align, int *start, int *plain, hipStream_t s);
hipError_t idx_scatter(const int *assign, int n, int *cursor, int *perm, hipStream_t s);
hipError_t idx_centroids(const float *pm, int pm_ld, int dim, const int *perm, const int *start,
                         const int *counts, int nb, float *cent, int cent_ld, hipStream_t s);
hipError_t idx_transpose(const float *src, int ld, int nrows, int dim, float *dst, hipStream_t s);
hipError_t idx_gather_pack(const float *pm_src, int pm_ld, int dim, int oct, const int *perm,
                           int n_pos, int n_pad, float *packed, float *pm_dst, float *norms,
                           hipStream_t s);
hipError_t idx_pick_blocks(const float *score, int Q, int nb, int P, const float *lb,
                           const int *counts, int *chosen, float *lb_rest, hipStream_t s);
hipError_t idx_router_first(const float *score, int Q, int nb, const int *counts, float *lb,
                            hipStream_t s);
// Units outside [0, nu) (and -1) are no unit.
hipError_t idx_pair_count(const int *chosen, int Q, int P, int p0, int p1, int nu, int *counts,
                          hipStream_t s);
hipError_t idx_pair_scatter(const int *chosen, int Q, int P, int p0, int p1, int nu, int *cursor,
                            int *qlist, int *rows, int row_base, hipStream_t s);
hipError_t idx_gather_columns(const float *src, long long ld, int dim, const int *list, int n,
                              float *dst, hipStream_t s);
hipError_t idx_scatter_results(const float *dsub, const int *isub, int k, int n, const int *list,
                               int Q, float *dist, int *idx, hipStream_t s);
// Per-unit list starts (qstart, nu + 1 entries; multiples of qpt, queries
// a task), the scatter cursors (= qstart) and the scan tasks (slots past the
// last task written empty) from the units' pair counts.  clear (nullable):
// a counter set to 0 in stream order (the failure count of the select that
// follows), saving a memset node.
hipError_t idx_unit_tasks(const int *counts, int nu, int qpt, const int *unit_start, const int *unit_count,
                          int *qstart, int *cursor, ScanTask *tasks, int max_tasks, int *clear, hipStream_t s);
hipError_t idx_ref_block(const int *block_start, const int *block_count, int nb, const int *perm,
                         int *out, hipStream_t s);
hipError_t idx_row_norms(const float *pm, int ld, int n, int dim, float *out, hipStream_t s);

hipError_t idx_kpp_seed(const float *pm, int ld, int dim, long long n, int m, int nb,
                        unsigned long long seed, float *mind, int *cidx, double *wsum,
                        unsigned *ticket, float *cent, hipStream_t s);
hipError_t idx_group_fine(const int *cand, int kc, int nf, const int *fcount, int nb, int cap,
                          int *fblock, hipStream_t s);
hipError_t idx_block_count(const int *fblock, const int *fcount, int nf, int *bcount, int *bfine,
                           hipStream_t s);
hipError_t idx_fine_start(const int *fblock, const int *fcount, int nf, int *pos_cursor,
                          int *fstart, int *slot_cursor, int *fine_sorted, hipStream_t s);
hipError_t idx_finalize_fine(const float *pm, int ld, int dim, const float *fcent,
                             const int *fine_sorted, const int *fstart, const int *fcount,
                             const int *fblock, int nf, float *cent_s, double *rad_s, int *block_s,
                             hipStream_t s);
hipError_t idx_ball_aux(const float *fnorm, const double *frad, int nf, float *fsq, float *fradf,
                        hipStream_t s);
// Ball probes (knn_index.hip): the nearest ball of each query's first
// block into out[q * Pb], then, from the threshold theta that scan left,
// up to cap more balls of its P blocks into out[q * Pb + 1 ..] with the
// others' lower bounds folded into lb_rest.  res: keys approximate squared
// distances (else ||r||^2 - 2 r.q with error coefficients eq, er).  At most
// idx_ball_cand_max() candidate balls per query.
hipError_t idx_ball_first(const float *qpm, int pm_ld, int dim, int Q, const int *chosen, int P,
                          const float *fcent, const double *frad, const int *fb_start, int *out, int Pb,
                          hipStream_t s);
hipError_t idx_ball_rest(const float *qpm, int pm_ld, int dim, int Q, const float *qnorm, const unsigned *theta,
                         const unsigned *rmax, double eq, double er, int res, const int *chosen, int P,
                         const float *fcent, const double *frad, const int *fb_start, int *out, int Pb, int cap,
                         float *lb_rest, hipStream_t s);
int idx_ball_cand_max();
// Every ball of a one-pass search at once: out[q * Pb ..] the balls of
// query q's P blocks whose bound is below the reach of the nearest balls
// holding kp members (fcnt: members per ball), nearest first, the others'
// bounds folded into lb_rest.  ncand: most candidate balls of a query (P
// times the most balls of a block), at most idx_ball_cand_max().
// theta (nullable): every query's scan threshold is written: seeded from
// T_q with at most 128 candidates (order-mapped keys ||r||^2 - 2 r.q with
// error coefficients eq, er; rmax: the largest ||r||^2), else +inf.
hipError_t idx_ball_select(const float *qpm, int pm_ld, int dim, int Q, int kp, const int *chosen, int P,
                           const float *fcent, const double *frad, const int *fcnt, const int *fb_start, int *out,
                           int Pb, int ncand, float *lb_rest, hipStream_t s, const float *qnorm = nullptr,
                           const unsigned *rmax = nullptr, double eq = 0.0, double er = 0.0,
                           unsigned *theta = nullptr);
// Radii of balls with known position ranges (largest member distance, FP64;
// -1 for an empty ball).
hipError_t idx_ball_range_radius(const float *pm, int ld, int dim, const float *fcent, const int *ball_pos,
                                 const int *ball_cnt, int nf, double *rad, hipStream_t s);
// Ball radii refitted to the block-ordered points pm (a loaded index):
// every point joins the nearest ball of its block.  d2bits / member: nf
// words of scratch each; rad: nf radii (-1: no member).
hipError_t idx_ball_refit(const float *pm, int ld, int dim, const int *block_start, const int *block_count,
                          const int *fb_start, int nb, const float *fcent, int nf,
                          unsigned long long *d2bits, int *member, double *rad, hipStream_t s);
// cbf (nullable): bfloat16 hi/lo ball centres (idx_pack_ball_bf3) for
// three-term bound products over nch_bf 16-dim chunks (dim <= 128), with
// error coefficient g15_bf; else FP32 products.
// The tiled bound matrix (lb_at) as nb rows of Q.
hipError_t idx_lb_untile(const float *tiled, int Q, int nb, float *out, hipStream_t s);
hipError_t idx_ball_lb(const float *qpm, int pm_ld, int Q, int dim, const float *qnorm,
                       const float *fcent, int nf, const float *fnorm, const float *fsq,
                       const float *fradf, const int *block_s, const int *fb_start, int nb, float *lb,
                       hipStream_t s, const unsigned short *cbf = nullptr, int nch_bf = 0,
                       float g15_bf = 0.0f);
// Slot form of the bounds (knn_index.hip, ball_lb_slot_kernel): each block
// in the smallest of 4, 8, 16 or 32 aligned slots that holds its balls, per
// slot the ball (sball, -1 empty), per tile a mode word and per group of 4
// slots the block it stores (built by the host from the blocks' ball counts).
// The same bounds as idx_ball_lb, every entry of the matrix written.
hipError_t idx_ball_slot_gather(const int *sball, int nsl, const float *fcent, int pm_ld, const float *fnorm,
                                const float *fsq, const float *fradf, float *scent, float *snrm, float *ssq,
                                float *srad, hipStream_t s);
bool idx_ball_lb_slots_fit(int dim, int pm_ld, int nch_bf);
hipError_t idx_ball_lb_slots(const float *qpm, int pm_ld, int Q, int dim, const float *qnorm, const float *scent,
                             const unsigned short *sbf, int nch_bf, float g15_bf, int nsl, const float *snrm,
                             const float *ssq, const float *srad, const int *sblk, const int *smode, int nb,
                             float *lb, hipStream_t s);
hipError_t idx_pack_ball_bf3(const float *fcent, int pm_ld, int nf, int dim, int nch, unsigned short *out,
                             hipStream_t s);
// Outputs of one problem of a batched class: its queries are columns
// qstart .. qstart + qn - 1 of the class's k x Qtot results.
struct BatchOut {
    float *dist;
    int *idx;
    int qn;
    int qstart;
};
hipError_t idx_batch_outputs(const float *dist, const int *idx, int k, int Qtot, const BatchOut *outs,
                             int np, hipStream_t s);
hipError_t idx_batch_perm(const int *starts, const int *counts, int nsets, int n_pos, int *perm,
                          hipStream_t s);

}  // namespace knnhip
