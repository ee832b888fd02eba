// Block index for the router form: k-means blocks over the reference set,
// references stored block-contiguously, and a search that scans each query's
// most promising blocks first and skips the others only when the triangle
// inequality proves they cannot hold a neighbour.
//
// The result stays the reference ground truth (code/test.cpp:129-175) bit
// for bit: the scanned candidates go through the same exact re-rank and
// certificate as the plain path, and the certificate also requires every
// unscanned block's lower bound max(0, |q - c_B| - rho_B)^2 to clear the k-th
// key; a query that fails is answered by the exact full scan.
#include <float.h>

#include "knn_common.h"
#include "knn_index.h"

namespace knnhip {

// ---------------------------------------------------------------------------
// k-means support
// ---------------------------------------------------------------------------

// counts[assign[i]]++ for i < n.
__global__ void histogram_kernel(const int *__restrict__ assign, int n, int *__restrict__ counts) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        atomicAdd(&counts[assign[i]], 1);
}

// Exclusive prefix sums of counts (one block; nb <= 65536): start[b] over
// counts rounded up to `align`, and plain[b] over the raw counts.
__global__ void prefix_kernel(const int *__restrict__ counts, int nb, int align,
                              int *__restrict__ start, int *__restrict__ plain) {
    __shared__ int part[1024];
    __shared__ int partp[1024];
    const int per = (nb + blockDim.x - 1) / blockDim.x;
    const int b0 = threadIdx.x * per;
    int s = 0, sp = 0;
    for (int b = b0; b < min(nb, b0 + per); ++b) {
        s += (counts[b] + align - 1) / align * align;
        sp += counts[b];
    }
    part[threadIdx.x] = s;
    partp[threadIdx.x] = sp;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0, accp = 0;
        for (int t = 0; t < (int)blockDim.x; ++t) {
            const int v = part[t], vp = partp[t];
            part[t] = acc;
            partp[t] = accp;
            acc += v;
            accp += vp;
        }
        start[nb] = acc;
        if (plain) plain[nb] = accp;
    }
    __syncthreads();
    s = part[threadIdx.x];
    sp = partp[threadIdx.x];
    for (int b = b0; b < min(nb, b0 + per); ++b) {
        start[b] = s;
        if (plain) plain[b] = sp;
        s += (counts[b] + align - 1) / align * align;
        sp += counts[b];
    }
}

// Scatter i into slot cursor[assign[i]]++ (order within a block is by
// arrival; the exactness argument does not depend on it).
__global__ void scatter_kernel(const int *__restrict__ assign, int n, int *__restrict__ cursor,
                               int *__restrict__ perm) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int slot = atomicAdd(&cursor[assign[i]], 1);
        perm[slot] = i;
    }
}

// Centroid of each block from its members (point-major source rows);
// blocks without members keep their previous centroid.  One block per
// k-means block, one thread per dimension chunk.
__global__ void centroid_kernel(const float *__restrict__ pm, int pm_ld, int dim,
                                const int *__restrict__ perm, const int *__restrict__ start,
                                const int *__restrict__ counts, float *__restrict__ cent,
                                int cent_ld) {
    const int b = blockIdx.x;
    const int n = counts[b];
    if (n == 0) return;
    const int s0 = start[b];
    for (int d = threadIdx.x; d < dim; d += blockDim.x) {
        double acc = 0.0;
        for (int j = 0; j < n; ++j) acc += (double)pm[(long long)perm[s0 + j] * pm_ld + d];
        cent[(long long)b * cent_ld + d] = (float)(acc / n);
    }
}

// Point-major [k][ld] -> dim-major [d][k] (the engine's input layout).
__global__ void transpose_kernel(const float *__restrict__ src, int ld, int nrows, int dim,
                                 float *__restrict__ dst) {
    const long long total = (long long)nrows * dim;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long d = e / nrows, i = e - d * nrows;
        dst[e] = src[i * ld + d];
    }
}

// Initial centroids: references at an even stride.
__global__ void init_centroids_kernel(const float *__restrict__ pm, int pm_ld, int n, int dim,
                                      int nb, float *__restrict__ cent, int cent_ld) {
    const long long total = (long long)nb * dim;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long b = e / dim, d = e - b * dim;
        const long long src = (b * (long long)n) / nb;
        cent[b * cent_ld + d] = pm[src * pm_ld + d];
    }
}

// Reordered layouts: position i of the block-contiguous order holds
// reference perm[i] (or nothing: perm[i] < 0 pads a block to 32).  Writes the
// packed tiles (as the norms pass does), the point-major copy and the norms.
__global__ void gather_pack_kernel(const float *__restrict__ pm_src, int pm_ld, int dim, int oct,
                                   const int *__restrict__ perm, int n_pos, int n_pad,
                                   float *__restrict__ packed, float *__restrict__ pm_dst,
                                   float *__restrict__ norms) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pad) return;
    const int src = i < n_pos ? perm[i] : -1;
    const float *row = src >= 0 ? pm_src + (long long)src * pm_ld : nullptr;
    const int tile = i / kRefsPerTile, m = i % kRefsPerTile;
    float s = 0.0f;
    for (int j = 0; j < oct; ++j) {
        float v[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int d = 8 * j + t;
            v[t] = (row && d < dim) ? row[d] : 0.0f;
            s = fmaf(v[t], v[t], s);
        }
        float *dst = packed + (((long long)tile * oct + j) * kWave + m) * 4;
        *reinterpret_cast<float4 *>(dst) = make_float4(v[0], v[1], v[2], v[3]);
        *reinterpret_cast<float4 *>(dst + 32 * 4) = make_float4(v[4], v[5], v[6], v[7]);
        if (i < n_pos) {
            float *out = pm_dst + (long long)i * pm_ld;
            if (8 * j + 4 <= pm_ld)
                *reinterpret_cast<float4 *>(out + 8 * j) = make_float4(v[0], v[1], v[2], v[3]);
            if (8 * j + 8 <= pm_ld)
                *reinterpret_cast<float4 *>(out + 8 * j + 4) = make_float4(v[4], v[5], v[6], v[7]);
        }
    }
    if (i < n_pos) norms[i] = s;
}

// Block radius rho_B = max over members of |r - c_B|, in double, then
// inflated by a relative 1e-12 so that it bounds the real distance.
__global__ void radius_kernel(const float *__restrict__ pm, int pm_ld, int dim,
                              const int *__restrict__ start, const int *__restrict__ counts,
                              const float *__restrict__ cent, int cent_ld,
                              double *__restrict__ radius) {
    const int b = blockIdx.x;
    const int n = counts[b];
    const int s0 = start[b];
    const float *c = cent + (long long)b * cent_ld;
    double best = 0.0;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const float *r = pm + (long long)(s0 + j) * pm_ld;
        double acc = 0.0;
        for (int d = 0; d < dim; ++d) {
            const double t = (double)r[d] - (double)c[d];
            acc += t * t;
        }
        best = fmax(best, acc);
    }
    __shared__ double red[256];
    red[threadIdx.x] = best;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) radius[b] = sqrt(red[0]) * (1.0 + 1e-12);
}

// ---------------------------------------------------------------------------
// Search support
// ---------------------------------------------------------------------------

// Lower bound of every reference of block B for query q, rounded down to
// float: max(0, |q - c_B| - rho_B), |q - c_B| in double.  One thread per
// (query, block); lb is [q][nb].
__global__ void lower_bound_kernel(const float *__restrict__ qpm, int pm_ld, int Q, int dim,
                                   const float *__restrict__ cent, int cent_ld, int nb,
                                   const double *__restrict__ radius, float *__restrict__ lb) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)Q * nb) return;
    const int q = (int)(e / nb), b = (int)(e - (long long)q * nb);
    const float *qr = qpm + (long long)q * pm_ld;
    const float *c = cent + (long long)b * cent_ld;
    double acc = 0.0;
    for (int d = 0; d < dim; ++d) {
        const double t = (double)qr[d] - (double)c[d];
        acc += t * t;
    }
    const double v = sqrt(acc) * (1.0 - 1e-12) - radius[b];
    float f = (float)v;
    if ((double)f > v) f = nextafterf(f, -INFINITY);
    lb[e] = f > 0.0f ? f : 0.0f;
}

// Per query: the P blocks of largest score (ties: lower block), and the
// smallest lower bound over the blocks not taken.  One wave per query.
// Blocks without members are never taken (they hold no references).
__global__ __launch_bounds__(256) void pick_blocks_kernel(const float *__restrict__ score, int Q,
                                                          int nb, int P,
                                                          const float *__restrict__ lb,
                                                          const int *__restrict__ counts,
                                                          int *__restrict__ chosen,
                                                          float *__restrict__ lb_rest) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int q = blockIdx.x * (blockDim.x >> 6) + wave;
    if (q >= Q) return;
    const float *sc = score + (long long)q * nb;
    const float *l = lb + (long long)q * nb;
    unsigned long long taken_mask_dummy = 0;
    (void)taken_mask_dummy;
    // Per round: wave argmax over blocks not yet taken.
    float last_s = INFINITY;
    int last_b = -1;
    for (int p = 0; p < P; ++p) {
        float bs = -INFINITY;
        int bb = 0x7FFFFFFF;
        for (int b = lane; b < nb; b += kWave) {
            if (counts[b] == 0) continue;
            const float s = sc[b];
            // strictly after the previous pick in (score desc, block asc) order
            const bool after = (s < last_s) || (s == last_s && b > last_b);
            if (!after) continue;
            if (s > bs || (s == bs && b < bb)) {
                bs = s;
                bb = b;
            }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const float os = __shfl_xor(bs, off);
            const int ob = __shfl_xor(bb, off);
            if (os > bs || (os == bs && ob < bb)) {
                bs = os;
                bb = ob;
            }
        }
        if (lane == 0) chosen[(long long)q * P + p] = bb == 0x7FFFFFFF ? -1 : bb;
        last_s = bs;
        last_b = bb;
    }
    // smallest lower bound over the blocks not taken (those after the last
    // pick in the same order), empty blocks excluded
    float m = INFINITY;
    for (int b = lane; b < nb; b += kWave) {
        if (counts[b] == 0) continue;
        const float s = sc[b];
        const bool after = (s < last_s) || (s == last_s && b > last_b);
        if (after) m = fminf(m, l[b]);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) m = fminf(m, __shfl_xor(m, off));
    if (lane == 0) lb_rest[q] = m;
}

// Invert the (query -> blocks) lists, probes [p0, p1) of each query's P, into
// per-block query lists and scan tasks of up to 32 queries.  Step 1: count
// pairs per block.
__global__ void pair_count_kernel(const int *__restrict__ chosen, int Q, int P, int p0, int p1,
                                  int *__restrict__ counts) {
    const int w = p1 - p0;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < (long long)Q * w;
         e += (long long)gridDim.x * blockDim.x) {
        const long long q = e / w;
        const int b = chosen[q * P + p0 + (e - q * w)];
        if (b >= 0) atomicAdd(&counts[b], 1);
    }
}

// Step 2: place each pair in its block's list (qlist, grouped by block with
// each block's group starting at a multiple of 32) and record the heap row
// it will produce: the pair's slot in the list, plus row_base, is its
// output row.
__global__ void pair_scatter_kernel(const int *__restrict__ chosen, int Q, int P, int p0, int p1,
                                    int *__restrict__ cursor, int *__restrict__ qlist,
                                    int *__restrict__ rows, int row_base) {
    const int w = p1 - p0;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < (long long)Q * w;
         e += (long long)gridDim.x * blockDim.x) {
        const long long q = e / w;
        const long long pi = q * P + p0 + (e - q * w);
        const int b = chosen[pi];
        if (b < 0) {
            rows[pi] = -1;
            continue;
        }
        const int slot = atomicAdd(&cursor[b], 1);
        qlist[slot] = (int)q;
        rows[pi] = row_base + slot;
    }
}

// Gather query columns list[0..n) of a dim-major set into a compact one,
// and scatter k x n results back into the k x Q outputs.
__global__ void gather_columns_kernel(const float *__restrict__ src, long long ld, int dim,
                                      const int *__restrict__ list, int n, float *__restrict__ dst) {
    const long long total = (long long)n * dim;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long d = e / n, j = e - d * n;
        dst[e] = src[d * ld + list[j]];
    }
}
__global__ void scatter_results_kernel(const float *__restrict__ dsub, const int *__restrict__ isub,
                                       int k, int n, const int *__restrict__ list, int Q,
                                       float *__restrict__ dist, int *__restrict__ idx) {
    const long long total = (long long)n * k;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long r = e / n, j = e - r * n;
        dist[r * Q + list[j]] = dsub[e];
        idx[r * Q + list[j]] = isub[e];
    }
}

// Step 3: tasks.  Block b's pairs occupy qlist[qstart[b] .. qstart[b] +
// count[b]) with qstart[b] a multiple of 32; task t covers list slots
// [32t, 32t+32) (slot == output row), so tasks are enumerated from the
// aligned slot space directly.
__global__ void make_tasks_kernel(const int *__restrict__ qstart, const int *__restrict__ counts,
                                  int nb, const int *__restrict__ block_start,
                                  const int *__restrict__ block_count, ScanTask *__restrict__ tasks,
                                  int max_tasks) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        const int c = counts[b];
        for (int j = 0; j < c; j += kQueriesPerWave) {
            const int t = (qstart[b] + j) / kQueriesPerWave;
            if (t >= max_tasks) continue;
            ScanTask tk;
            tk.ref_begin = block_start[b];
            tk.ref_end = block_start[b] + block_count[b];
            tk.qoff = qstart[b] + j;
            tk.qcount = min(kQueriesPerWave, c - j);
            tasks[t] = tk;
        }
    }
}

// Empty task slots (alignment gaps between blocks) scan nothing.
__global__ void clear_tasks_kernel(ScanTask *__restrict__ tasks, int n) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        ScanTask tk;
        tk.ref_begin = 0;
        tk.ref_end = 0;
        tk.qoff = 0;
        tk.qcount = 0;
        tasks[t] = tk;
    }
}

// Negated lower bound as the ranking score of the router-free search
// (nearest block first).
__global__ void negate_kernel(const float *__restrict__ x, long long n, float *__restrict__ y) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x)
        y[e] = -x[e];
}

// ---------------------------------------------------------------------------
// Two-level layout: fine clusters (sub-balls) grouped into blocks
// ---------------------------------------------------------------------------

// bcount[fblock[f]] += fcount[f]
__global__ void block_count_kernel(const int *__restrict__ fblock, const int *__restrict__ fcount,
                                   int nf, int *__restrict__ bcount) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x)
        atomicAdd(&bcount[fblock[f]], fcount[f]);
}

// Each fine cluster takes a contiguous range inside its block.
__global__ void fine_start_kernel(const int *__restrict__ fblock, const int *__restrict__ fcount,
                                  int nf, int *__restrict__ bcursor, int *__restrict__ fstart) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x)
        fstart[f] = atomicAdd(&bcursor[fblock[f]], fcount[f]);
}

// Coarse centroid = mean of its fine centroids weighted by their sizes (a
// block with no members keeps its centroid).  One block per coarse cluster;
// members listed by fine_by_block[fb_start[b] .. fb_start[b+1]).
__global__ void coarse_centroid_kernel(const float *__restrict__ fcent, int ld, int dim,
                                       const int *__restrict__ fine_by_block,
                                       const int *__restrict__ fb_start,
                                       const int *__restrict__ fcount, float *__restrict__ cent) {
    const int b = blockIdx.x;
    const int f0 = fb_start[b], f1 = fb_start[b + 1];
    long long w = 0;
    for (int j = f0; j < f1; ++j) w += fcount[fine_by_block[j]];
    if (w == 0) return;
    for (int d = threadIdx.x; d < dim; d += blockDim.x) {
        double acc = 0.0;
        for (int j = f0; j < f1; ++j) {
            const int f = fine_by_block[j];
            acc += (double)fcent[(long long)f * ld + d] * (double)fcount[f];
        }
        cent[(long long)b * ld + d] = (float)(acc / (double)w);
    }
}

// Lower bound of every reference of fine cluster f for query q (distance,
// not squared, rounded down): max(0, |q - c_f| - rho_f).  |q - c_f|^2 comes
// from ||q||^2 + ||c||^2 - 2 q.c on FP32 MFMA, lowered by the error bound
// g*(||q||^2 + ||c||^2 + 2|q||c|), g = 2*(dim+4)*2^-24.  One wave computes a
// 32 (clusters) x 32 (queries) tile; k-step kk reads dimension kk + h*KH on
// lane half h (KH = half the padded dimension), 32 k-steps per chunk.
typedef float f32x16_t __attribute__((ext_vector_type(16)));

__global__ __launch_bounds__(256) void fine_lb_kernel(const float *__restrict__ qpm, int pm_ld, int Q,
                                                      int dim, const float *__restrict__ qnorm,
                                                      const float *__restrict__ fcent, int nf,
                                                      const float *__restrict__ fnorm,
                                                      const double *__restrict__ frad,
                                                      float *__restrict__ lbf) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 31, h = lane >> 5;
    const int ntf = (nf + 31) / 32;
    const long long tile = (long long)blockIdx.x * (blockDim.x >> 6) + wave;
    const int tq = (int)(tile / ntf), tf = (int)(tile - (long long)tq * ntf);
    const int q0 = tq * 32, f0 = tf * 32;
    if (q0 >= Q) return;
    const int qa = min(q0 + col, Q - 1);  // B operand: query col
    const int fa = min(f0 + col, nf - 1); // A operand: cluster col
    const float *qr = qpm + (long long)qa * pm_ld;
    const float *cr = fcent + (long long)fa * pm_ld;
    const int KH = (dim + 1) / 2;
    f32x16_t acc = {};
    if ((KH & 3) == 0 && (pm_ld & 3) == 0) {
        // each lane reads its own row 16 bytes at a time (dimensions kk..kk+3 of its half)
        const float *ca = cr + h * KH, *qa4 = qr + h * KH;
        for (int k0 = 0; k0 < KH; k0 += 16) {
            float4 av[4], bv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const bool ok = k0 + 4 * u < KH;
                av[u] = ok ? *reinterpret_cast<const float4 *>(ca + k0 + 4 * u) : make_float4(0, 0, 0, 0);
                bv[u] = ok ? *reinterpret_cast<const float4 *>(qa4 + k0 + 4 * u) : make_float4(0, 0, 0, 0);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[u].x, bv[u].x, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[u].y, bv[u].y, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[u].z, bv[u].z, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[u].w, bv[u].w, acc, 0, 0, 0);
            }
        }
    } else {
        for (int k0 = 0; k0 < KH; k0 += 32) {
            float av[32], bv[32];
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const int kk = k0 + u;
                const int d = kk + h * KH;
                const bool ok = kk < KH && d < dim;
                av[u] = ok ? cr[d] : 0.0f;
                bv[u] = ok ? qr[d] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < 32; ++u)
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[u], bv[u], acc, 0, 0, 0);
        }
    }
    const int q = q0 + col;
    if (q >= Q) return;
    const double g = 2.0 * (double)(dim + 4) * 5.9604644775390625e-8;
    const double nq = (double)qnorm[q];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int f = f0 + ((i >> 2) << 3) + (h << 2) + (i & 3);
        if (f >= nf) continue;
        const double nc = (double)fnorm[f];
        const double dot = (double)acc[i];
        double d2 = nq + nc - 2.0 * dot - g * (nq + nc + 2.0 * sqrt(nq * nc)) * 1.5;
        d2 = d2 > 0.0 ? d2 : 0.0;
        const double v = sqrt(d2) * (1.0 - 1e-12) - frad[f];
        float fv = (float)v;
        if ((double)fv > v) fv = nextafterf(fv, -INFINITY);
        lbf[(long long)q * nf + f] = fv > 0.0f ? fv : 0.0f;
    }
}

// Squared norms of point-major rows (FP32, sequential).
__global__ void row_norms_kernel(const float *__restrict__ pm, int ld, int n, int dim,
                                 float *__restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int d = 0; d < dim; ++d) s = fmaf(pm[(long long)i * ld + d], pm[(long long)i * ld + d], s);
        out[i] = s;
    }
}

// Block lower bound = min over the block's fine clusters; negated as the
// router-free ranking score.  lbb and score are [q][nb].
__global__ void block_lb_kernel(const float *__restrict__ lbf, int Q, int nf,
                                const int *__restrict__ fine_by_block,
                                const int *__restrict__ fb_start, int nb, float *__restrict__ lbb,
                                float *__restrict__ score) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)Q * nb) return;
    const int q = (int)(e / nb), b = (int)(e - (long long)q * nb);
    float m = INFINITY;
    for (int j = fb_start[b]; j < fb_start[b + 1]; ++j)
        m = fminf(m, lbf[(long long)q * nf + fine_by_block[j]]);
    lbb[e] = m;
    if (score) score[e] = -m;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static int grid_for(long long n, int threads = 256, int cap = 65536) {
    long long g = (n + threads - 1) / threads;
    return (int)(g < 1 ? 1 : (g > cap ? cap : g));
}

hipError_t idx_histogram(const int *assign, int n, int *counts, hipStream_t s) {
    hipLaunchKernelGGL(histogram_kernel, dim3(grid_for(n)), dim3(256), 0, s, assign, n, counts);
    return hipGetLastError();
}
hipError_t idx_prefix(const int *counts, int nb, int align, int *start, int *plain, hipStream_t s) {
    hipLaunchKernelGGL(prefix_kernel, dim3(1), dim3(1024), 0, s, counts, nb, align, start, plain);
    return hipGetLastError();
}
hipError_t idx_scatter(const int *assign, int n, int *cursor, int *perm, hipStream_t s) {
    hipLaunchKernelGGL(scatter_kernel, dim3(grid_for(n)), dim3(256), 0, s, assign, n, cursor, perm);
    return hipGetLastError();
}
hipError_t idx_centroids(const float *pm, int pm_ld, int dim, const int *perm, const int *start,
                         const int *counts, int nb, float *cent, int cent_ld, hipStream_t s) {
    hipLaunchKernelGGL(centroid_kernel, dim3(nb), dim3(128), 0, s, pm, pm_ld, dim, perm, start,
                       counts, cent, cent_ld);
    return hipGetLastError();
}
hipError_t idx_transpose(const float *src, int ld, int nrows, int dim, float *dst, hipStream_t s) {
    hipLaunchKernelGGL(transpose_kernel, dim3(grid_for((long long)nrows * dim)), dim3(256), 0, s,
                       src, ld, nrows, dim, dst);
    return hipGetLastError();
}
hipError_t idx_init_centroids(const float *pm, int pm_ld, int n, int dim, int nb, float *cent,
                              int cent_ld, hipStream_t s) {
    hipLaunchKernelGGL(init_centroids_kernel, dim3(grid_for((long long)nb * dim)), dim3(256), 0, s,
                       pm, pm_ld, n, dim, nb, cent, cent_ld);
    return hipGetLastError();
}
hipError_t idx_gather_pack(const float *pm_src, int pm_ld, int dim, int oct, const int *perm,
                           int n_pos, int n_pad, float *packed, float *pm_dst, float *norms,
                           hipStream_t s) {
    hipLaunchKernelGGL(gather_pack_kernel, dim3((n_pad + 255) / 256), dim3(256), 0, s, pm_src,
                       pm_ld, dim, oct, perm, n_pos, n_pad, packed, pm_dst, norms);
    return hipGetLastError();
}
hipError_t idx_radius(const float *pm, int pm_ld, int dim, const int *start, const int *counts,
                      int nb, const float *cent, int cent_ld, double *radius, hipStream_t s) {
    hipLaunchKernelGGL(radius_kernel, dim3(nb), dim3(256), 0, s, pm, pm_ld, dim, start, counts,
                       cent, cent_ld, radius);
    return hipGetLastError();
}
hipError_t idx_lower_bounds(const float *qpm, int pm_ld, int Q, int dim, const float *cent,
                            int cent_ld, int nb, const double *radius, float *lb, hipStream_t s) {
    const long long n = (long long)Q * nb;
    hipLaunchKernelGGL(lower_bound_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, qpm,
                       pm_ld, Q, dim, cent, cent_ld, nb, radius, lb);
    return hipGetLastError();
}
hipError_t idx_pick_blocks(const float *score, int Q, int nb, int P, const float *lb,
                           const int *counts, int *chosen, float *lb_rest, hipStream_t s) {
    hipLaunchKernelGGL(pick_blocks_kernel, dim3((Q + 3) / 4), dim3(256), 0, s, score, Q, nb, P, lb,
                       counts, chosen, lb_rest);
    return hipGetLastError();
}
hipError_t idx_pair_count(const int *chosen, int Q, int P, int p0, int p1, int *counts,
                          hipStream_t s) {
    hipLaunchKernelGGL(pair_count_kernel, dim3(grid_for((long long)Q * (p1 - p0))), dim3(256), 0, s,
                       chosen, Q, P, p0, p1, counts);
    return hipGetLastError();
}
hipError_t idx_pair_scatter(const int *chosen, int Q, int P, int p0, int p1, int *cursor,
                            int *qlist, int *rows, int row_base, hipStream_t s) {
    hipLaunchKernelGGL(pair_scatter_kernel, dim3(grid_for((long long)Q * (p1 - p0))), dim3(256), 0,
                       s, chosen, Q, P, p0, p1, cursor, qlist, rows, row_base);
    return hipGetLastError();
}
hipError_t idx_gather_columns(const float *src, long long ld, int dim, const int *list, int n,
                              float *dst, hipStream_t s) {
    hipLaunchKernelGGL(gather_columns_kernel, dim3(grid_for((long long)n * dim)), dim3(256), 0, s,
                       src, ld, dim, list, n, dst);
    return hipGetLastError();
}
hipError_t idx_scatter_results(const float *dsub, const int *isub, int k, int n, const int *list,
                               int Q, float *dist, int *idx, hipStream_t s) {
    hipLaunchKernelGGL(scatter_results_kernel, dim3(grid_for((long long)n * k)), dim3(256), 0, s,
                       dsub, isub, k, n, list, Q, dist, idx);
    return hipGetLastError();
}
hipError_t idx_make_tasks(const int *qstart, const int *counts, int nb, const int *block_start,
                          const int *block_count, ScanTask *tasks, int max_tasks, hipStream_t s) {
    hipLaunchKernelGGL(clear_tasks_kernel, dim3(grid_for(max_tasks)), dim3(256), 0, s, tasks,
                       max_tasks);
    hipLaunchKernelGGL(make_tasks_kernel, dim3(grid_for(nb)), dim3(256), 0, s, qstart, counts, nb,
                       block_start, block_count, tasks, max_tasks);
    return hipGetLastError();
}
hipError_t idx_negate(const float *x, long long n, float *y, hipStream_t s) {
    hipLaunchKernelGGL(negate_kernel, dim3(grid_for(n)), dim3(256), 0, s, x, n, y);
    return hipGetLastError();
}

hipError_t idx_block_count(const int *fblock, const int *fcount, int nf, int *bcount,
                           hipStream_t s) {
    hipLaunchKernelGGL(block_count_kernel, dim3(grid_for(nf)), dim3(256), 0, s, fblock, fcount, nf,
                       bcount);
    return hipGetLastError();
}
hipError_t idx_fine_start(const int *fblock, const int *fcount, int nf, int *bcursor, int *fstart,
                          hipStream_t s) {
    hipLaunchKernelGGL(fine_start_kernel, dim3(grid_for(nf)), dim3(256), 0, s, fblock, fcount, nf,
                       bcursor, fstart);
    return hipGetLastError();
}
hipError_t idx_coarse_centroids(const float *fcent, int ld, int dim, const int *fine_by_block,
                                const int *fb_start, const int *fcount, int nb, float *cent,
                                hipStream_t s) {
    hipLaunchKernelGGL(coarse_centroid_kernel, dim3(nb), dim3(128), 0, s, fcent, ld, dim,
                       fine_by_block, fb_start, fcount, cent);
    return hipGetLastError();
}
hipError_t idx_fine_lb(const float *qpm, int pm_ld, int Q, int dim, const float *qnorm,
                       const float *fcent, int nf, const float *fnorm, const double *frad,
                       float *lbf, hipStream_t s) {
    const long long tiles = (long long)((Q + 31) / 32) * ((nf + 31) / 32);
    hipLaunchKernelGGL(fine_lb_kernel, dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, s, qpm, pm_ld,
                       Q, dim, qnorm, fcent, nf, fnorm, frad, lbf);
    return hipGetLastError();
}
hipError_t idx_row_norms(const float *pm, int ld, int n, int dim, float *out, hipStream_t s) {
    hipLaunchKernelGGL(row_norms_kernel, dim3(grid_for(n)), dim3(256), 0, s, pm, ld, n, dim, out);
    return hipGetLastError();
}
hipError_t idx_block_lb(const float *lbf, int Q, int nf, const int *fine_by_block,
                        const int *fb_start, int nb, float *lbb, float *score, hipStream_t s) {
    const long long n = (long long)Q * nb;
    hipLaunchKernelGGL(block_lb_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, lbf, Q,
                       nf, fine_by_block, fb_start, nb, lbb, score);
    return hipGetLastError();
}

}  // namespace knnhip
