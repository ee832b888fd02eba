// Learned router for the block index: an MLP  dim -> H -> H -> nblocks
// (ReLU after the two hidden layers) whose logits rank the blocks of each
// query.  Trained on the GPU with the kernels below (BF16 MFMA products with
// FP32 accumulation, softmax cross-entropy against the blocks that hold each
// training query's exact neighbours, Adam).  The router only orders blocks:
// the index's certificate decides what may be skipped, so results stay
// exact whatever the network outputs.
//
// Numerics (mirrored by the oracle's restatement): every product operand is
// rounded to bfloat16 (round to nearest even) when it is read; products
// accumulate in FP32; bias, ReLU, softmax and Adam run in FP32.
#include <math.h>

#include "knn_common.h"
#include "knn_router.h"

namespace knnhip {

typedef short s16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

__device__ inline unsigned short to_bf16(float f) {
    const unsigned u = __float_as_uint(f);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (unsigned short)((u >> 16) | 0x40u);  // NaN stays NaN
    return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// ---------------------------------------------------------------------------
// C[M][N] = op(A) op(B) (+ bias) (ReLU) (* [mask > 0]), FP32 in and out.
// op(A)(m,k) = A[k*lda + m] if TA else A[m*lda + k];
// op(B)(k,n) = B[n*ldb + k] if TB else B[k*ldb + n].
// A workgroup (4 waves, 2 x 2) computes a 64 x 64 tile; each K chunk of 32
// is staged in LDS as bfloat16, rows indexed by m (A) and n (B) with k
// contiguous, so an MFMA operand (4 consecutive k of one row) is one 8-byte
// LDS read.  v_mfma_f32_32x32x8_bf16: lane l supplies row l%32, k = 4(l/32)
// .. +3 of the 8-deep step; result register i is row 8(i/4) + 4(l/32) + i%4,
// column l%32.
// ---------------------------------------------------------------------------
constexpr int kTile = 64, kKc = 32, kLdsPitch = kKc + 8;  // bf16 elements per LDS row

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) unsigned short As[kTile * kLdsPitch];
    __shared__ __attribute__((aligned(16))) unsigned short Bs[kTile * kLdsPitch];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
    f32x16_t acc = {};
    for (int k0 = 0; k0 < g.K; k0 += kKc) {
#pragma unroll
        for (int r = 0; r < (kTile * kKc) / 256; ++r) {
            const int e = t + 256 * r;
            int mm, kk;
            if (TA) {
                mm = e % kTile;
                kk = e / kTile;
            } else {
                mm = e / kKc;
                kk = e % kKc;
            }
            const int m = m0 + mm, k = k0 + kk;
            float v = 0.0f;
            if (m < g.M && k < g.K) v = TA ? g.A[(long long)k * g.lda + m] : g.A[(long long)m * g.lda + k];
            As[mm * kLdsPitch + kk] = to_bf16(v);
        }
#pragma unroll
        for (int r = 0; r < (kTile * kKc) / 256; ++r) {
            const int e = t + 256 * r;
            int nn, kk;
            if (TB) {
                nn = e / kKc;
                kk = e % kKc;
            } else {
                nn = e % kTile;
                kk = e / kTile;
            }
            const int n = n0 + nn, k = k0 + kk;
            float v = 0.0f;
            if (n < g.N && k < g.K) v = TB ? g.B[(long long)n * g.ldb + k] : g.B[(long long)k * g.ldb + n];
            Bs[nn * kLdsPitch + kk] = to_bf16(v);
        }
        __syncthreads();
#pragma unroll
        for (int k8 = 0; k8 < kKc; k8 += 8) {
            const int ko = k8 + 4 * (lane >> 5);
            const s16x4_t a =
                *reinterpret_cast<const s16x4_t *>(&As[(wm * 32 + (lane & 31)) * kLdsPitch + ko]);
            const s16x4_t b =
                *reinterpret_cast<const s16x4_t *>(&Bs[(wn * 32 + (lane & 31)) * kLdsPitch + ko]);
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const int n = n0 + wn * 32 + (lane & 31);
    if (n >= g.N) return;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int m = m0 + wm * 32 + ((i >> 2) << 3) + ((lane >> 5) << 2) + (i & 3);
        if (m >= g.M) continue;
        float v = acc[i];
        if (g.bias_mode == 1) v += g.bias[n];
        if (g.bias_mode == 2) v += g.bias[m];
        if (g.relu) v = fmaxf(v, 0.0f);
        if (g.mask && !(g.mask[(long long)m * g.ldm + n] > 0.0f)) v = 0.0f;
        g.C[(long long)m * g.ldc + n] = v;
    }
}

hipError_t launch_gemm(const GemmArgs &g, hipStream_t s) {
    const dim3 grid((g.N + kTile - 1) / kTile, (g.M + kTile - 1) / kTile);
    if (g.ta && g.tb)
        hipLaunchKernelGGL((gemm_bf16_kernel<true, true>), grid, dim3(256), 0, s, g);
    else if (g.ta)
        hipLaunchKernelGGL((gemm_bf16_kernel<true, false>), grid, dim3(256), 0, s, g);
    else if (g.tb)
        hipLaunchKernelGGL((gemm_bf16_kernel<false, true>), grid, dim3(256), 0, s, g);
    else
        hipLaunchKernelGGL((gemm_bf16_kernel<false, false>), grid, dim3(256), 0, s, g);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Loss: softmax over the nb logits of each sample i (row of Z), target
// distribution t_b = (number of the sample's k neighbours in block b) / k,
// loss_i = -sum_b t_b log p_b, dZ = (p - t) * scale.  tgt[j*ldt + i] is the
// block of neighbour j of sample i.  One workgroup per sample.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void softmax_ce_kernel(const float *__restrict__ Z, int nb,
                                                         const int *__restrict__ tgt, long long ldt,
                                                         int k, float scale, float *__restrict__ dZ,
                                                         float *__restrict__ loss) {
    __shared__ float red[256];
    __shared__ int tb[kRouterMaxK];
    const int i = blockIdx.x, t = threadIdx.x;
    const float *z = Z + (long long)i * nb;
    for (int j = t; j < k; j += blockDim.x) tb[j] = tgt[(long long)j * ldt + i];
    float mx = -INFINITY;
    for (int b = t; b < nb; b += blockDim.x) mx = fmaxf(mx, z[b]);
    red[t] = mx;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w) red[t] = fmaxf(red[t], red[t + w]);
        __syncthreads();
    }
    mx = red[0];
    __syncthreads();
    float se = 0.0f;
    for (int b = t; b < nb; b += blockDim.x) se += expf(z[b] - mx);
    red[t] = se;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    se = red[0];
    const float lse = logf(se);
    const float inv_k = 1.0f / (float)k;
    for (int b = t; b < nb; b += blockDim.x) {
        int c = 0;
        for (int j = 0; j < k; ++j) c += tb[j] == b;
        const float p = expf(z[b] - mx) / se;
        dZ[(long long)i * nb + b] = (p - (float)c * inv_k) * scale;
    }
    if (t == 0) {
        float l = 0.0f;
        for (int j = 0; j < k; ++j) l -= (z[tb[j]] - mx - lse) * inv_k;
        loss[i] = l;
    }
}

// out[j] = sum_i X[i*ld + j] (FP32): bias gradients.  A workgroup takes 64
// columns; wave w sums rows w, w + kColsumWaves, ... in order and the
// partial sums are added in wave order, so the result is deterministic.
constexpr int kColsumWaves = 16;

__global__ __launch_bounds__(64 * kColsumWaves) void colsum_kernel(const float *__restrict__ X, int n,
                                                                   int m, long long ld,
                                                                   float *__restrict__ out) {
    __shared__ float part[kColsumWaves][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int j = blockIdx.x * 64 + lane;
    float s = 0.0f;
    if (j < m)
        for (int i = wave; i < n; i += kColsumWaves) s += X[(long long)i * ld + j];
    part[wave][lane] = s;
    __syncthreads();
    if (wave == 0 && j < m) {
        float t = part[0][lane];
        for (int w = 1; w < kColsumWaves; ++w) t += part[w][lane];
        out[j] = t;
    }
}

// Adam (Kingma & Ba 2015), the usual bias-corrected form:
// m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
// w -= lr * (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps).
__global__ void adam_kernel(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, long long n, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const float ge = g[e];
        const float me = b1 * m[e] + (1.0f - b1) * ge;
        const float ve = b2 * v[e] + (1.0f - b2) * ge * ge;
        m[e] = me;
        v[e] = ve;
        w[e] -= lr * (me / c1) / (sqrtf(ve / c2) + eps);
    }
}

// Uniform init in [-a_l, a_l] from a counter hash of (seed, element), with
// a_l = sqrt(6 / fan_in) for weights (He-uniform) and 0 for biases.
// Segment offsets, lengths and fan-ins come in `lay`.
__global__ void init_params_kernel(float *__restrict__ w, RouterLayout lay, unsigned long long seed) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < lay.total;
         e += (long long)gridDim.x * blockDim.x) {
        float a = 0.0f;
        for (int l = 0; l < 3; ++l)
            if (e >= lay.w_off[l] && e < lay.w_off[l] + lay.w_len[l]) a = sqrtf(6.0f / (float)lay.fan_in[l]);
        unsigned long long x = seed ^ ((unsigned long long)e * 0x9E3779B97F4A7C15ull);
        x += 0x9E3779B97F4A7C15ull;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        x ^= x >> 31;
        const float u = (float)(x >> 40) * 5.9604644775390625e-8f;  // [0, 1)
        w[e] = a * (2.0f * u - 1.0f);
    }
}

// tgt[j][i] = ref_block[nbr[j][i]] (blocks of each training query's k
// exact neighbours; both k x n).
__global__ void neighbour_blocks_kernel(const int *__restrict__ nbr, long long n, int k,
                                        const int *__restrict__ ref_block, int *__restrict__ tgt) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n * k;
         e += (long long)gridDim.x * blockDim.x)
        tgt[e] = ref_block[nbr[e]];
}

// dst[j][c] = src[j][list[c]] for j < rows, c < n (int columns).
__global__ void gather_int_columns_kernel(const int *__restrict__ src, long long ld, int rows,
                                          const int *__restrict__ list, int n, int *__restrict__ dst) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < (long long)rows * n;
         e += (long long)gridDim.x * blockDim.x) {
        const long long j = e / n, c = e - j * n;
        dst[e] = src[j * ld + list[c]];
    }
}

// Sum of n floats into *out (one workgroup).
__global__ void sum_kernel(const float *__restrict__ x, int n, float *__restrict__ out) {
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += x[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

static int grid_cap(long long n) {
    long long g = (n + 255) / 256;
    return (int)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

hipError_t launch_softmax_ce(const float *Z, int n, int nb, const int *tgt, long long ldt, int k,
                             float scale, float *dZ, float *loss, hipStream_t s) {
    hipLaunchKernelGGL(softmax_ce_kernel, dim3(n), dim3(256), 0, s, Z, nb, tgt, ldt, k, scale, dZ,
                       loss);
    return hipGetLastError();
}
hipError_t launch_colsum(const float *X, int n, int m, long long ld, float *out, hipStream_t s) {
    hipLaunchKernelGGL(colsum_kernel, dim3((m + 63) / 64), dim3(64 * kColsumWaves), 0, s, X, n, m, ld,
                       out);
    return hipGetLastError();
}
hipError_t launch_adam(float *w, const float *g, float *m, float *v, long long n, float lr, float b1,
                       float b2, float eps, int step, hipStream_t s) {
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(grid_cap(n)), dim3(256), 0, s, w, g, m, v, n, lr, b1, b2,
                       eps, c1, c2);
    return hipGetLastError();
}
hipError_t launch_init_params(float *w, const RouterLayout &lay, unsigned long long seed,
                              hipStream_t s) {
    hipLaunchKernelGGL(init_params_kernel, dim3(grid_cap(lay.total)), dim3(256), 0, s, w, lay, seed);
    return hipGetLastError();
}
hipError_t launch_neighbour_blocks(const int *nbr, long long n, int k, const int *ref_block,
                                   int *tgt, hipStream_t s) {
    hipLaunchKernelGGL(neighbour_blocks_kernel, dim3(grid_cap(n * k)), dim3(256), 0, s, nbr, n, k,
                       ref_block, tgt);
    return hipGetLastError();
}
hipError_t launch_gather_int_columns(const int *src, long long ld, int rows, const int *list, int n,
                                     int *dst, hipStream_t s) {
    hipLaunchKernelGGL(gather_int_columns_kernel, dim3(grid_cap((long long)rows * n)), dim3(256), 0, s,
                       src, ld, rows, list, n, dst);
    return hipGetLastError();
}
hipError_t launch_sum(const float *x, int n, float *out, hipStream_t s) {
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, s, x, n, out);
    return hipGetLastError();
}


// ---------------------------------------------------------------------------
// Inference: the first probe of each query is the non-empty block with the
// highest router score.  A bfloat16 copy of the weights (rows padded to 16
// elements, zero-filled) feeds one fused kernel per 64 queries: the three
// layers run back to back with the activations in LDS (bfloat16, as the
// training forward rounds them), and the last layer keeps a running argmax
// per query instead of writing the nb x Q score matrix.  Ties and NaN follow
// router_first_kernel: NaN scores count as -inf, equal scores keep the lower
// block.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));

__global__ void router_bf16_weights_kernel(const float *__restrict__ w, RouterLayout lay, int dim,
                                           int dpad, int H, int nb, unsigned short *__restrict__ out) {
    const long long n1 = (long long)H * dpad, n2 = (long long)H * H, n3 = (long long)nb * H;
    const long long n = n1 + n2 + n3;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        float v;
        if (e < n1) {
            const long long r = e / dpad, c = e - r * dpad;
            v = c < dim ? w[lay.w_off[0] + r * dim + c] : 0.0f;
        } else if (e < n1 + n2) {
            v = w[lay.w_off[1] + (e - n1)];
        } else {
            v = w[lay.w_off[2] + (e - n1 - n2)];
        }
        out[e] = to_bf16(v);
    }
}

constexpr int kRtQueries = 64;  // queries per workgroup (two 32-column tiles)
constexpr int kRtWaves = 8;

__device__ inline bool rt_better(float s, int b, float bs, int bb) {
    return bb < 0 || s > bs || (s == bs && b < bb);
}

// out[n][m] = bf16(relu(sum_k W[m][k] in[n][k] + bias[m])) for the block's
// 64 queries n and H units m; in / out are LDS rows of pitch ip / op.
template <int KH>
__device__ inline void rt_layer(const unsigned short *__restrict__ W, int wp, int ksteps,
                                const float *__restrict__ bias, const unsigned short *in, int ip,
                                unsigned short *out, int op, int wave, int lane) {
    constexpr int H = 16 * KH;
    const int h = lane >> 5, col = lane & 31;
    for (int t = wave; t < (H / 32) * 2; t += kRtWaves) {
        const int ut = t >> 1, qt = t & 1;
        const unsigned short *wr = W + (long long)(ut * 32 + col) * wp + 8 * h;
        const unsigned short *xr = in + (qt * 32 + col) * ip + 8 * h;
        f32x16_t acc = {};
        for (int ks = 0; ks < ksteps; ++ks) {
            const u32x4_t av = *reinterpret_cast<const u32x4_t *>(wr + 16 * ks);
            const u32x4_t bv = *reinterpret_cast<const u32x4_t *>(xr + 16 * ks);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, av),
                                                          __builtin_bit_cast(bf16x8_t, bv), acc, 0, 0, 0);
        }
        // element i: unit ut*32 + 8(i/4) + 4h + i%4, query qt*32 + col
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int u0 = ut * 32 + 8 * g + 4 * h;
            unsigned short v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) v[j] = to_bf16(fmaxf(acc[4 * g + j] + bias[u0 + j], 0.0f));
            uint2 pk;
            pk.x = (unsigned)v[0] | ((unsigned)v[1] << 16);
            pk.y = (unsigned)v[2] | ((unsigned)v[3] << 16);
            *reinterpret_cast<uint2 *>(out + (qt * 32 + col) * op + u0) = pk;
        }
    }
}

template <int KH>
__global__ __launch_bounds__(64 * kRtWaves) void router_first_fused_kernel(RouterFirstArgs a) {
    constexpr int H = 16 * KH, hp = H + 8;
    extern __shared__ __attribute__((aligned(16))) unsigned char rt_smem[];
    const int xp = a.dpad + 8;
    unsigned short *xs = reinterpret_cast<unsigned short *>(rt_smem);
    unsigned short *h1 = xs + kRtQueries * xp;
    unsigned short *h2 = h1 + kRtQueries * hp;
    float *cs = reinterpret_cast<float *>(h2 + kRtQueries * hp);
    int *cb = reinterpret_cast<int *>(cs + kRtWaves * 2 * kRtQueries);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
    const int q0 = blockIdx.x * kRtQueries;
    // queries to LDS as bfloat16 rows (coalesced over queries)
    for (int e = threadIdx.x; e < kRtQueries * a.dpad; e += 64 * kRtWaves) {
        const int j = e % kRtQueries, d = e / kRtQueries;
        const float v = (q0 + j < a.Q && d < a.dim) ? a.x[(long long)d * a.ldx + q0 + j] : 0.0f;
        xs[j * xp + d] = to_bf16(v);
    }
    __syncthreads();
    rt_layer<KH>(a.w1, a.dpad, a.dpad / 16, a.b1, xs, xp, h1, hp, wave, lane);
    __syncthreads();
    rt_layer<KH>(a.w2, H, KH, a.b2, h1, hp, h2, hp, wave, lane);
    __syncthreads();
    // last layer: the block's two query tiles as B operands, block tiles
    // w, w + 4, ... as A; running best per (lane, query tile)
    u32x4_t bq[2][KH];
#pragma unroll
    for (int qt = 0; qt < 2; ++qt)
#pragma unroll
        for (int ks = 0; ks < KH; ++ks)
            bq[qt][ks] = *reinterpret_cast<const u32x4_t *>(h2 + (qt * 32 + col) * hp + 16 * ks + 8 * h);
    float bs[2] = {-INFINITY, -INFINITY};
    int bb[2] = {-1, -1};
    const int ntb = (a.nb + 31) / 32;
    for (int bt = wave; bt < ntb; bt += kRtWaves) {
        const int row = min(bt * 32 + col, a.nb - 1);
        const unsigned short *wr = a.w3 + (long long)row * H + 8 * h;
        // the rows' biases and point counts, loaded with the tile's weights
        // (read after the products: their latency overlaps the MFMAs)
        float b3r[16];
        bool emp[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int b = min(bt * 32 + 8 * (i >> 2) + 4 * h + (i & 3), a.nb - 1);
            b3r[i] = a.b3[b];
            emp[i] = a.counts && a.counts[b] == 0;
        }
        f32x16_t acc0 = {}, acc1 = {};
#pragma unroll
        for (int ks = 0; ks < KH; ++ks) {
            const bf16x8_t av = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u32x4_t *>(wr + 16 * ks));
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, __builtin_bit_cast(bf16x8_t, bq[0][ks]), acc0,
                                                           0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, __builtin_bit_cast(bf16x8_t, bq[1][ks]), acc1,
                                                           0, 0, 0);
        }
        // element i: block bt*32 + 8(i/4) + 4h + i%4 (increasing in i)
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int b = bt * 32 + 8 * (i >> 2) + 4 * h + (i & 3);
            if (b >= a.nb || emp[i]) continue;
            const float b3 = b3r[i];
            float s0 = acc0[i] + b3, s1 = acc1[i] + b3;
            s0 = s0 == s0 ? s0 : -INFINITY;
            s1 = s1 == s1 ? s1 : -INFINITY;
            if (rt_better(s0, b, bs[0], bb[0])) {
                bs[0] = s0;
                bb[0] = b;
            }
            if (rt_better(s1, b, bs[1], bb[1])) {
                bs[1] = s1;
                bb[1] = b;
            }
        }
    }
#pragma unroll
    for (int qt = 0; qt < 2; ++qt) {
        cs[(wave * 2 + h) * kRtQueries + qt * 32 + col] = bs[qt];
        cb[(wave * 2 + h) * kRtQueries + qt * 32 + col] = bb[qt];
    }
    __syncthreads();
    if (threadIdx.x < kRtQueries && q0 + (int)threadIdx.x < a.Q) {
        const int j = threadIdx.x, q = q0 + j;
        float s = -INFINITY;
        int b = -1;
        for (int c = 0; c < kRtWaves * 2; ++c) {
            const int cbb = cb[c * kRtQueries + j];
            if (cbb >= 0 && rt_better(cs[c * kRtQueries + j], cbb, s, b)) {
                s = cs[c * kRtQueries + j];
                b = cbb;
            }
        }
        if (a.top1) a.top1[q] = b;
        if (a.lb && b >= 0) a.lb[lb_at(b, q, a.nb)] = -INFINITY;
    }
}

size_t router_first_lds(int dpad, int H) {
    return (size_t)kRtQueries * (dpad + 8) * 2 + 2 * (size_t)kRtQueries * (H + 8) * 2 +
           (size_t)kRtWaves * 2 * kRtQueries * 8;
}

bool router_first_supported(int dpad, int H) {
    return (H == 64 || H == 128 || H == 256) && dpad <= 512 && router_first_lds(dpad, H) <= 160 * 1024;
}

hipError_t launch_router_bf16_weights(const float *w, const RouterLayout &lay, int dim, int dpad, int H,
                                      int nb, unsigned short *out, hipStream_t s) {
    const long long n = (long long)H * dpad + (long long)H * H + (long long)nb * H;
    hipLaunchKernelGGL(router_bf16_weights_kernel, dim3(grid_cap(n)), dim3(256), 0, s, w, lay, dim, dpad,
                       H, nb, out);
    return hipGetLastError();
}

hipError_t launch_router_first_fused(const RouterFirstArgs &a, int H, hipStream_t s) {
    const size_t lds = router_first_lds(a.dpad, H);
    const dim3 grid((a.Q + kRtQueries - 1) / kRtQueries);
#define RT(KHV)                                                                                   \
    if (H == 16 * KHV) {                                                                          \
        if (lds > 65536) {                                                                        \
            const hipError_t e = hipFuncSetAttribute((const void *)router_first_fused_kernel<KHV>, \
                                                     hipFuncAttributeMaxDynamicSharedMemorySize,   \
                                                     (int)lds);                                    \
            if (e != hipSuccess) return e;                                                        \
        }                                                                                         \
        hipLaunchKernelGGL(router_first_fused_kernel<KHV>, grid, dim3(64 * kRtWaves), lds, s, a);  \
        return hipGetLastError();                                                                 \
    }
    RT(4)
    RT(8)
    RT(16)
#undef RT
    return hipErrorInvalidValue;
}

}  // namespace knnhip
