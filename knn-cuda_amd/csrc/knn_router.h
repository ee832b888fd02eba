// Kernels of the learned router (knn_router.hip).
#pragma once

#include "knn_common.h"

namespace knnhip {

// Largest neighbour count a training target may hold.
constexpr int kRouterMaxK = 256;

// C[M][N] = op(A) op(B) (+ bias) (ReLU) (* [mask > 0]); see knn_router.hip.
struct GemmArgs {
    const float *A;
    long long lda;
    bool ta;
    const float *B;
    long long ldb;
    bool tb;
    float *C;
    long long ldc;
    int M, N, K;
    const float *bias;
    int bias_mode;  // 0 none, 1 bias[n], 2 bias[m]
    bool relu;
    const float *mask;  // nullptr: none
    long long ldm;
};

// Flat parameter vector: W1 [H][dim], b1 [H], W2 [H][H], b2 [H], W3 [nb][H],
// b3 [nb] (FP32).
struct RouterLayout {
    long long w_off[3], w_len[3], b_off[3];
    int fan_in[3];
    long long total;
};

// Fused inference (first probe per query): queries x (dim-major, ldx), the
// bfloat16 weight copy (W1 rows padded to dpad, W2 and W3 rows of H), FP32
// biases; counts (nullable) marks empty blocks; outputs top1[q] (nullable)
// and bound (top1, q) = -inf of the tiled bound matrix lb (lb_at; nullable).
struct RouterFirstArgs {
    const float *x;
    long long ldx;
    int Q, dim, dpad;
    const unsigned short *w1, *w2, *w3;
    const float *b1, *b2, *b3;
    int nb;
    const int *counts;
    int *top1;
    float *lb;
};
bool router_first_supported(int dpad, int H);
hipError_t launch_router_bf16_weights(const float *w, const RouterLayout &lay, int dim, int dpad, int H,
                                      int nb, unsigned short *out, hipStream_t s);
hipError_t launch_router_first_fused(const RouterFirstArgs &a, int H, hipStream_t s);

hipError_t launch_gemm(const GemmArgs &g, hipStream_t s);
hipError_t launch_softmax_ce(const float *Z, int n, int nb, const int *tgt, long long ldt, int k,
                             float scale, float *dZ, float *loss, hipStream_t s);
hipError_t launch_colsum(const float *X, int n, int m, long long ld, float *out, hipStream_t s);
hipError_t launch_adam(float *w, const float *g, float *m, float *v, long long n, float lr, float b1,
                       float b2, float eps, int step, hipStream_t s);
hipError_t launch_init_params(float *w, const RouterLayout &lay, unsigned long long seed,
                              hipStream_t s);
hipError_t launch_neighbour_blocks(const int *nbr, long long n, int k, const int *ref_block,
                                   int *tgt, hipStream_t s);
hipError_t launch_gather_int_columns(const int *src, long long ld, int rows, const int *list, int n,
                                     int *dst, hipStream_t s);
hipError_t launch_sum(const float *x, int n, float *out, hipStream_t s);

}  // namespace knnhip
