// Block index for the router form: k-means blocks over the reference set,
// references stored block-contiguously, and a search that scans each query's
// most promising blocks first and skips the others only when the triangle
// inequality proves they cannot hold a neighbour.
//
// The result stays the reference ground truth (code/test.cpp:129-175) bit
// for bit: the scanned candidates go through the same exact re-rank and
// certificate as the plain path, and the certificate also requires every
// unscanned block's lower bound (the smallest max(0, |q - c_f| - rho_f)^2
// over the sub-balls f covering the block) to clear the k-th key; a query
// that fails is answered by the plain full scan.
#include <float.h>

#include <algorithm>

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

// Exclusive prefix sums of counts (one block of up to 1024 threads):
// start[b] over counts rounded up to `align`, and plain[b] over the raw
// counts, with the totals at [nb].  The block takes blockDim.x counts at a
// time: a scan within each wave (lane shifts), the earlier waves' totals
// from LDS, and a carry across the chunks.
__global__ void prefix_kernel(const int *__restrict__ counts, int nb, int align,
                              int *__restrict__ start, int *__restrict__ plain) {
    __shared__ int wsum[16], wsump[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    // the first chunks' counts are all read before the scans start
    constexpr int kAhead = 8;
    int cnts[kAhead];
#pragma unroll
    for (int i = 0; i < kAhead; ++i) {
        const int b = i * blockDim.x + threadIdx.x;
        cnts[i] = b < nb ? counts[b] : 0;
    }
    int base = 0, basep = 0;
    for (int c0 = 0, ci = 0; c0 < nb; c0 += blockDim.x, ++ci) {
        const int b = c0 + threadIdx.x;
        int cnt = 0;
#pragma unroll
        for (int i = 0; i < kAhead; ++i)
            if (i == ci) cnt = cnts[i];
        if (ci >= kAhead) cnt = b < nb ? counts[b] : 0;
        const int own = (cnt + align - 1) / align * align;
        int v = own, vp = cnt;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const int x = __shfl_up(v, off), xp = __shfl_up(vp, off);
            if (lane >= off) {
                v += x;
                vp += xp;
            }
        }
        if (lane == kWave - 1) {
            wsum[wave] = v;
            wsump[wave] = vp;
        }
        __syncthreads();
        int before = 0, beforep = 0, tot = 0, totp = 0;
        for (int w = 0; w < nw; ++w) {
            if (w < wave) {
                before += wsum[w];
                beforep += wsump[w];
            }
            tot += wsum[w];
            totp += wsump[w];
        }
        if (b < nb) {
            start[b] = base + before + v - own;
            if (plain) plain[b] = basep + beforep + vp - cnt;
        }
        base += tot;
        basep += totp;
        __syncthreads();  // wsum is written again by the next chunk
    }
    if (threadIdx.x == 0) {
        start[nb] = base;
        if (plain) plain[nb] = basep;
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

// ---------------------------------------------------------------------------
// Search support
// ---------------------------------------------------------------------------

// Per query: the P blocks of largest score (ties: lower block), and the
// smallest lower bound over the blocks not taken.  score and lb are
// block-major ([b][q], query_nb per row); score == nullptr ranks by -lb
// (nearest block first); a NaN score ranks last.  A workgroup serves 64
// queries (one per lane, so reads are whole rows) with kPickWaves waves,
// each over its share of the blocks; every round each wave offers its best
// block strictly after the previous pick and wave 0 keeps the best offer.
// Blocks without members are never taken (they hold no references).
constexpr int kPickWaves = 16;
// Block field of the pick keys (blocks below 2^30).
constexpr int kPickKeyB = 0x3FFFFFFF;
// At most this many queries, one workgroup per query instead.
constexpr int kPickSingleMaxQ = 1024;

__device__ inline float rank_score(const float *__restrict__ score, const float *__restrict__ lb, int b,
                                   int q, int Q, int nb) {
    const float s = score ? score[(long long)b * Q + q] : -lb[lb_at(b, q, nb)];
    return s == s ? s : -INFINITY;
}

__global__ __launch_bounds__(64 * kPickWaves) void pick_blocks_kernel(
    const float *__restrict__ score, int Q, int nb, int P, const float *__restrict__ lb,
    const int *__restrict__ counts, int *__restrict__ chosen, float *__restrict__ lb_rest) {
    __shared__ float offer_s[kPickWaves][64];
    __shared__ int offer_b[kPickWaves][64];
    __shared__ float keep_s[64];
    __shared__ int keep_b[64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int q = blockIdx.x * 64 + lane;
    const bool valid = q < Q;
    const int per = (nb + kPickWaves - 1) / kPickWaves;
    const int b0 = min(nb, wave * per), b1 = min(nb, b0 + per);
    // (last_s, last_b): the previous pick; "after" = lower score, or equal
    // score and higher block
    float last_s = INFINITY;
    int last_b = -1;
    for (int p = 0; p < P; ++p) {
        float bs = -INFINITY;
        int bb = -1;
        if (valid)
            for (int b = b0; b < b1; ++b) {
                if (counts[b] == 0) continue;
                const float s = rank_score(score, lb, b, q, Q, nb);
                const bool after = (s < last_s) || (s == last_s && b > last_b);
                if (after && (bb < 0 || s > bs)) {
                    bs = s;
                    bb = b;
                }
            }
        offer_s[wave][lane] = bs;
        offer_b[wave][lane] = bb;
        __syncthreads();
        if (wave == 0) {
            float ks = -INFINITY;
            int kb = -1;
            for (int w = 0; w < kPickWaves; ++w) {
                const int ob = offer_b[w][lane];
                const float os = offer_s[w][lane];
                if (ob >= 0 && (kb < 0 || os > ks || (os == ks && ob < kb))) {
                    ks = os;
                    kb = ob;
                }
            }
            if (valid) chosen[(long long)q * P + p] = kb;
            // nothing left: no block is after (-inf, INT_MAX)
            keep_s[lane] = kb >= 0 ? ks : -INFINITY;
            keep_b[lane] = kb >= 0 ? kb : 0x7FFFFFFF;
        }
        __syncthreads();
        last_s = keep_s[lane];
        last_b = keep_b[lane];
    }
    // smallest lower bound over the blocks after the last pick
    float m = INFINITY;
    if (valid)
        for (int b = b0; b < b1; ++b) {
            if (counts[b] == 0) continue;
            const float s = rank_score(score, lb, b, q, Q, nb);
            const bool after = (s < last_s) || (s == last_s && b > last_b);
            if (after) m = fminf(m, lb[lb_at(b, q, nb)]);
        }
    offer_s[wave][lane] = m;
    __syncthreads();
    if (wave == 0 && valid) {
        for (int w = 1; w < kPickWaves; ++w) m = fminf(m, offer_s[w][lane]);
        lb_rest[q] = m;
    }
}

// Single-pass form for P <= PMAX: every thread keeps the best P blocks of its
// share (sorted, in registers) and the smallest bound over the rest of its
// share; P merge rounds then take the best head among the 16 waves' lists.
// lb_rest = min(bounds a thread did not keep, kept entries not chosen).
// Same result as pick_blocks_kernel, reading the bounds once.
template <int PMAX, bool SC>
__global__ __launch_bounds__(64 * kPickWaves) void pick_blocks_list_kernel(
    const float *__restrict__ score, int Q, int nb, int P, const float *__restrict__ lb,
    const int *__restrict__ counts, int *__restrict__ chosen, float *__restrict__ lb_rest) {
    __shared__ float offer_s[kPickWaves][64];
    __shared__ int offer_b[kPickWaves][64];
    __shared__ int winner[64];
    // (the wave number read as uniform: the block loop's bounds and addresses stay scalar)
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int q = blockIdx.x * 64 + lane;
    const bool valid = q < Q;
    const int per = (nb + kPickWaves - 1) / kPickWaves;
    const int b0 = min(nb, wave * per), b1 = min(nb, b0 + per);
    // the kept blocks as 64-bit keys, sorted: the rank score order-mapped
    // (-0 taken as +0, NaN as -inf), then the complemented block index, so a
    // larger key ranks first and equal scores go by the smaller block; 0
    // marks an unfilled slot.  SC (scores given): kl, their bounds; else a
    // bound is the negated score.
    unsigned long long key[SC ? PMAX : 1];
    double dkey[SC ? 1 : PMAX];
    double rkey = 0.0;  // (no score form) the largest key that left the list
    float kl[SC ? PMAX : 1];
    float rest = INFINITY;
#pragma unroll
    for (int j = 0; j < (SC ? PMAX : 1); ++j) key[j] = 0;
#pragma unroll
    for (int j = 0; j < (SC ? 1 : PMAX); ++j) dkey[j] = j < P ? 0.0 : (double)INFINITY;  // (+inf: a slot past P, passed by)
#pragma unroll
    for (int j = 0; j < (SC ? PMAX : 1); ++j) kl[j] = INFINITY;
    // bounds are read kPickBatch blocks at a time (independent loads, the
    // next batch in flight while one is taken, PF), then taken in block order
    constexpr bool PF = !(SC && PMAX > 8);  // (sixteen slots with scores: no room for two batches)
    constexpr int kPickBatch = PF ? 8 : 4;
    // (the loop is uniform over the wave, so the blocks' point counts are
    // scalar loads; lanes past the last query read the last one's bounds)
    const int qa = valid ? q : Q - 1;
    // (the query's column of its slab: block b's bound at lbq[64 b])
    const float *lbq = lb + lb_at(0, qa, nb);
    float lv[kPickBatch], sv[SC ? kPickBatch : 1];
    auto load = [&](int bb) {
#pragma unroll
        for (int j = 0; j < kPickBatch; ++j) {
            const int b = min(bb + j, b1 - 1);
            lv[j] = lbq[64 * b];
            if constexpr (SC) sv[j] = score[(long long)b * Q + qa];
        }
    };
    if (PF && b0 < b1) load(b0);
    for (int bb = b0; bb < b1; bb += kPickBatch) {
        if (!PF) load(bb);
        float cs_[kPickBatch], cl_[SC ? kPickBatch : 1];
        int cc_[kPickBatch];  // the blocks' point counts (0: empty, skipped)
#pragma unroll
        for (int j = 0; j < kPickBatch; ++j) {
            cc_[j] = counts[min(bb + j, b1 - 1)];
            if constexpr (SC) {
                cs_[j] = sv[j];
                cl_[j] = lv[j];
            } else {
                cs_[j] = -lv[j];
            }
        }
        if (PF && bb + kPickBatch < b1) load(bb + kPickBatch);
#pragma unroll
        for (int j = 0; j < kPickBatch; ++j) {
            const int b = bb + j;
            // (uniform: the block is the same for every lane)
            const bool live = b < b1 && cc_[j] != 0;
            if constexpr (!SC) {
                // the key as a positive double: bits 30..61 the order-mapped
                // score, bits 0..29 kPickKeyB - b; every such key is a normal
                // double (the order-mapped -inf is 0x007FFFFF, so the exponent
                // field is at least 1) and 0 marks an unfilled slot, so double
                // max/min order them as the integers: the new block takes its
                // place by a max/min pass and whatever leaves the last slot
                // (the smallest of P + 1) is folded into the rest key.  A block
                // past the share or without members enters as key 0, which
                // passes every slot and leaves the rest key as it is (no
                // branch, so the max/min results stay known canonical)
                const float sc = cs_[j] == cs_[j] ? cs_[j] + 0.0f : -INFINITY;
                const unsigned long long ck = ((unsigned long long)float_to_ordered(sc) << 30) | (unsigned)(kPickKeyB - b);
                double c = __longlong_as_double((long long)(live ? ck : 0ull));
#pragma unroll
                for (int t = 0; t < PMAX; ++t) {
                    const double kt = dkey[t];
                    dkey[t] = __builtin_fmax(kt, c);
                    c = __builtin_fmin(kt, c);
                }
                rkey = __builtin_fmax(rkey, c);
            } else {
                if (!live) continue;
                // one compare-swap pass over the P kept slots: the new block takes
                // the first unfilled slot or the first slot of a smaller key and
                // pushes the rest down; whatever leaves the last slot, the new
                // block itself or a kept one, is not chosen and bounds the rest
                const float sc = cs_[j] == cs_[j] ? cs_[j] + 0.0f : -INFINITY;
                unsigned long long c = ((unsigned long long)float_to_ordered(sc) << 32) | (unsigned)(~b);
                float cl = SC ? cl_[SC ? j : 0] : 0.0f;
#pragma unroll
                for (int t = 0; t < PMAX; ++t) {
                    if (t < P) {
                        const bool sw = c > key[t];
                        const unsigned long long tk = key[t];
                        key[t] = sw ? c : tk;
                        c = sw ? tk : c;
                        if constexpr (SC) {
                            const float tl = kl[t];
                            kl[t] = sw ? cl : tl;
                            cl = sw ? tl : cl;
                        }
                    }
                }
                // (an unfilled slot leaves key 0: nothing)
                const float out_lb = SC ? cl : c ? -ordered_to_float((unsigned)(c >> 32)) : INFINITY;
                rest = fminf(rest, out_lb);
            }
        }
    }
    if constexpr (!SC) {
        const unsigned long long rk = (unsigned long long)__double_as_longlong(rkey);
        if (rk) rest = -ordered_to_float((unsigned)(rk >> 30));
    }
    float ls[PMAX], ll[PMAX];
    int lbk[PMAX];
    int n = 0;
#pragma unroll
    for (int j = 0; j < PMAX; ++j) {
        bool real;
        if constexpr (SC) {
            real = key[SC ? j : 0] != 0;
            ls[j] = real ? ordered_to_float((unsigned)(key[SC ? j : 0] >> 32)) : -INFINITY;
            lbk[j] = real ? (int)~(unsigned)key[SC ? j : 0] : 0x7FFFFFFF;
        } else {
            const unsigned long long dk = (unsigned long long)__double_as_longlong(dkey[SC ? 0 : j]);
            real = j < P && dk != 0;
            ls[j] = real ? ordered_to_float((unsigned)(dk >> 30)) : -INFINITY;
            lbk[j] = real ? kPickKeyB - (int)(dk & kPickKeyB) : 0x7FFFFFFF;
        }
        ll[j] = SC ? kl[SC ? j : 0] : real ? -ls[j] : INFINITY;
        n += real ? 1 : 0;
    }
    // merge rounds
    int head = 0;
    for (int p = 0; p < P; ++p) {
        float hs = -INFINITY;
        int hb = -1;
#pragma unroll
        for (int j = 0; j < PMAX; ++j)
            if (j == head && j < n) {
                hs = ls[j];
                hb = lbk[j];
            }
        offer_s[wave][lane] = hs;
        offer_b[wave][lane] = hb;
        __syncthreads();
        if (wave == 0) {
            float ks = -INFINITY;
            int kb = -1, kw = -1;
            for (int w = 0; w < kPickWaves; ++w) {
                const int ob = offer_b[w][lane];
                const float os = offer_s[w][lane];
                if (ob >= 0 && (kb < 0 || os > ks || (os == ks && ob < kb))) {
                    ks = os;
                    kb = ob;
                    kw = w;
                }
            }
            if (valid) chosen[(long long)q * P + p] = kb;
            winner[lane] = kw;
        }
        __syncthreads();
        if (winner[lane] == wave) ++head;
        __syncthreads();
    }
    // entries of this thread's list that were not chosen
#pragma unroll
    for (int j = 0; j < PMAX; ++j)
        if (j >= head && j < n) rest = fminf(rest, ll[j]);
    offer_s[wave][lane] = rest;
    __syncthreads();
    if (wave == 0 && valid) {
        float m = offer_s[0][lane];
        for (int w = 1; w < kPickWaves; ++w) m = fminf(m, offer_s[w][lane]);
        lb_rest[q] = m;
    }
}

// The same selection for few queries: one workgroup per query, its threads
// spread over the blocks, so each round is a workgroup argmax.  The query's
// rank scores are staged in LDS first (NaN marks an empty block; rank_score
// never returns NaN), so the P rounds read LDS only.
__global__ __launch_bounds__(256) void pick_blocks_single_kernel(
    const float *__restrict__ score, int Q, int nb, int P, const float *__restrict__ lb,
    const int *__restrict__ counts, int *__restrict__ chosen, float *__restrict__ lb_rest) {
    extern __shared__ float sc[];  // nb rank scores
    __shared__ float red_s[4];
    __shared__ int red_b[4];
    const int q = blockIdx.x, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (int b = tid; b < nb; b += blockDim.x)
        sc[b] = counts[b] == 0 ? NAN : rank_score(score, lb, b, q, Q, nb);
    __syncthreads();
    float last_s = INFINITY;
    int last_b = -1;
    for (int p = 0; p < P; ++p) {
        float bs = -INFINITY;
        int bb = -1;
        for (int b = tid; b < nb; b += blockDim.x) {
            const float s = sc[b];
            if (s != s) continue;
            const bool after = (s < last_s) || (s == last_s && b > last_b);
            if (after && (bb < 0 || s > bs)) {
                bs = s;
                bb = b;
            }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const float os = __shfl_xor(bs, off);
            const int ob = __shfl_xor(bb, off);
            if (ob >= 0 && (bb < 0 || os > bs || (os == bs && ob < bb))) {
                bs = os;
                bb = ob;
            }
        }
        if (lane == 0) {
            red_s[wave] = bs;
            red_b[wave] = bb;
        }
        __syncthreads();
        float ks = -INFINITY;
        int kb = -1;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            const int ob = red_b[w];
            const float os = red_s[w];
            if (ob >= 0 && (kb < 0 || os > ks || (os == ks && ob < kb))) {
                ks = os;
                kb = ob;
            }
        }
        __syncthreads();
        if (tid == 0) chosen[(long long)q * P + p] = kb;
        last_s = kb >= 0 ? ks : -INFINITY;
        last_b = kb >= 0 ? kb : 0x7FFFFFFF;
    }
    float m = INFINITY;
    for (int b = tid; b < nb; b += blockDim.x) {
        const float s = sc[b];
        if (s != s) continue;
        const bool after = (s < last_s) || (s == last_s && b > last_b);
        if (after) m = fminf(m, lb[lb_at(b, q, nb)]);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) m = fminf(m, __shfl_xor(m, off));
    if (lane == 0) red_s[wave] = m;
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fminf(m, red_s[w]);
        lb_rest[q] = m;
    }
}

// Router-first ranking: each query's highest-scored non-empty block (ties to
// the lower block) gets the bound -inf, so that the bound ranking takes it
// as the first probe and ranks the others by their bounds.  A workgroup
// takes 64 queries (one per lane) and splits the blocks over its 16 waves
// (scores read kBatch blocks at a time); the waves' best are merged in LDS.
__global__ __launch_bounds__(64 * kPickWaves) void router_first_kernel(
    const float *__restrict__ score, int Q, int nb, const int *__restrict__ counts,
    float *__restrict__ lb) {
    constexpr int kBatch = 16;
    __shared__ float best_s[kPickWaves][64];
    __shared__ int best_b[kPickWaves][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int q = blockIdx.x * 64 + lane;
    const int per = (nb + kPickWaves - 1) / kPickWaves;
    const int b0 = min(nb, wave * per), b1 = min(nb, b0 + per);
    float bs = -INFINITY;
    int bb = -1;
    for (int bq = b0; q < Q && bq < b1; bq += kBatch) {
        float sv[kBatch];
        bool live[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int b = min(bq + j, b1 - 1);
            live[j] = bq + j < b1 && counts[b] != 0;
            sv[j] = rank_score(score, lb, b, q, Q, nb);
        }
#pragma unroll
        for (int j = 0; j < kBatch; ++j)
            if (live[j] && (bb < 0 || sv[j] > bs)) {
                bs = sv[j];
                bb = bq + j;
            }
    }
    best_s[wave][lane] = bs;
    best_b[wave][lane] = bb;
    __syncthreads();
    if (wave == 0 && q < Q) {
        // waves hold increasing block ranges: strict > keeps the lower block
        for (int w = 1; w < kPickWaves; ++w)
            if (best_b[w][lane] >= 0 && (bb < 0 || best_s[w][lane] > bs)) {
                bs = best_s[w][lane];
                bb = best_b[w][lane];
            }
        if (bb >= 0) lb[lb_at(bb, q, nb)] = -INFINITY;
    }
}

// ---------------------------------------------------------------------------
// Ball probes.  Inside the blocks a query probes, the scan can visit balls
// (each a contiguous position range of its block) instead of whole blocks:
// first the ball of the query's first block whose centre is nearest, then,
// once that scan has set the query's threshold, every other ball of its
// probed blocks whose lower bound lies below the distance the threshold
// stands for (nearest bounds first, at most `cap`).  Each ball left out
// lowers the query's bound on unscanned references to its own lower bound,
// so the threshold only steers the work: exactness rests on the certificate.
// ---------------------------------------------------------------------------

// Squared FP64 distances from a query row to the centre rows f[j] of up to
// kDistGroup balls (f[j] < 0: none), wave-uniform f, the whole wave
// cooperating: lane l takes dimensions 4l .. 4l + 3, then 4l + 256 .., so the
// loads of one centre row are one coalesced read and the group's rows are in
// flight together.  Each difference is exact in FP64 and each fused
// square-add rounds once: every sum is within (dim + 8) * 2^-53 relative of
// the true squared distance, whatever the order.  Every lane returns every
// sum.  Rows are 16-byte aligned; dimensions at or past dim are not read
// into the sums.
constexpr int kDistGroup = 8;
__device__ inline void wave_dist2_group(const float *__restrict__ q, const float *__restrict__ cent, int ld,
                                        int dim, const int (&f)[kDistGroup], double (&out)[kDistGroup],
                                        int lane) {
#pragma unroll
    for (int j = 0; j < kDistGroup; ++j) out[j] = 0.0;
    for (int d = 4 * lane; d < dim; d += 4 * kWave) {
        const float4 x = *reinterpret_cast<const float4 *>(q + d);
        float4 y[kDistGroup];
#pragma unroll
        for (int j = 0; j < kDistGroup; ++j)
            y[j] = f[j] >= 0 ? *reinterpret_cast<const float4 *>(cent + (long long)f[j] * ld + d) : x;
        const bool w1 = d + 1 < dim, w2 = d + 2 < dim, w3 = d + 3 < dim;
#pragma unroll
        for (int j = 0; j < kDistGroup; ++j) {
            double t = (double)x.x - (double)y[j].x;
            out[j] = fma(t, t, out[j]);
            t = w1 ? (double)x.y - (double)y[j].y : 0.0;
            out[j] = fma(t, t, out[j]);
            t = w2 ? (double)x.z - (double)y[j].z : 0.0;
            out[j] = fma(t, t, out[j]);
            t = w3 ? (double)x.w - (double)y[j].w : 0.0;
            out[j] = fma(t, t, out[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < kDistGroup; ++j)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) out[j] += __shfl_xor(out[j], o);
}

// Lower bound, rounded down to FP32, on the distance from a query to every
// member of a ball: FP64 centre distance d (relative error far below 1e-12)
// less the ball's FP64 radius.
__device__ inline float ball_member_lb(double d, double rho) {
    return __double2float_rd(d * (1.0 - 1e-12) - rho);
}

constexpr int kBallWaves = 4;  // queries (one wave each) per workgroup of the ball kernels

// First ball: out[q * Pb] = the non-empty ball of block chosen[q * P] whose
// centre is nearest to query q (ties: lower ball), or -1.
__global__ __launch_bounds__(64 * kBallWaves) void ball_first_kernel(
    const float *__restrict__ qpm, int pm_ld, int dim, int Q, const int *__restrict__ chosen, int P,
    const float *__restrict__ fcent, const double *__restrict__ frad, const int *__restrict__ fb_start,
    int *__restrict__ out, int Pb) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int q = blockIdx.x * kBallWaves + wave;
    if (q >= Q) return;
    const int b = chosen[(long long)q * P];
    double best = INFINITY;
    int bf = -1;
    if (b >= 0) {
        const float *qr = qpm + (long long)q * pm_ld;
        const int fl = fb_start[b], fh = fb_start[b + 1];
        for (int base = fl; base < fh; base += kDistGroup) {
            int f[kDistGroup];
#pragma unroll
            for (int j = 0; j < kDistGroup; ++j) f[j] = base + j < fh && frad[base + j] >= 0.0 ? base + j : -1;
            double d2[kDistGroup];
            wave_dist2_group(qr, fcent, pm_ld, dim, f, d2, lane);
            // ascending balls: a strict test keeps the lower one on ties
#pragma unroll
            for (int j = 0; j < kDistGroup; ++j)
                if (f[j] >= 0 && d2[j] < best) {
                    best = d2[j];
                    bf = f[j];
                }
        }
    }
    if (lane == 0) out[(long long)q * Pb] = bf;
}

// Other balls: for query q, every non-empty ball of blocks chosen[q * P ..
// q * P + P) except out[q * Pb] is a candidate with bound lb_f.  Those with
// lb_f below T_q, nearest bound first (ties: lower ball), fill out[q * Pb +
// 1 ..] up to cap entries (the rest -1); every other candidate lowers
// lb_rest[q] to lb_f.  T_q: with keys ||r||^2 - 2 r.q (res == 0),
// sqrt(theta + ||q||^2 (1 + g + eq) + er max||r||^2), the distance the
// threshold theta stands for widened by the key error bound; with keys that
// approximate squared distances (res != 0), sqrt(theta) * 1.02 + 1e-3.
// Candidates per query: at most kBallCand (the host checks P times the
// largest ball count of a block against it).
constexpr int kBallCand = 512;
__global__ __launch_bounds__(64 * kBallWaves) void ball_rest_kernel(
    const float *__restrict__ qpm, int pm_ld, int dim, int Q, const float *__restrict__ qnorm,
    const unsigned *__restrict__ theta, const unsigned *__restrict__ rmax, double eq, double er, int res,
    const int *__restrict__ chosen, int P, const float *__restrict__ fcent, const double *__restrict__ frad,
    const int *__restrict__ fb_start, int *__restrict__ out, int Pb, int cap, float *__restrict__ lb_rest) {
    __shared__ float c_lb[kBallWaves][kBallCand];
    __shared__ int c_f[kBallWaves][kBallCand];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int q = blockIdx.x * kBallWaves + wave;
    if (q >= Q) return;
    float *lbs = c_lb[wave];
    int *fs = c_f[wave];
    const long long o0 = (long long)q * Pb;
    const int f0 = out[o0];
    const float *qr = qpm + (long long)q * pm_ld;
    // candidates, block by block: slot n + (f - first ball of the block)
    int n = 0;
    for (int j = 0; j < P; ++j) {
        const int b = chosen[(long long)q * P + j];
        if (b < 0) continue;
        const int fl = fb_start[b], fh = fb_start[b + 1];
        for (int f = fl + lane; f < fh; f += kWave) fs[n + f - fl] = f != f0 && frad[f] >= 0.0 ? f : -1;
        n += fh - fl;
    }
    wave_sync();
    // their bounds, kDistGroup at a time (lane j keeps candidate base + j's)
    for (int base = 0; base < n; base += kDistGroup) {
        int f[kDistGroup];
#pragma unroll
        for (int j = 0; j < kDistGroup; ++j) f[j] = base + j < n ? fs[base + j] : -1;
        double d2[kDistGroup];
        wave_dist2_group(qr, fcent, pm_ld, dim, f, d2, lane);
        float lb = INFINITY;
#pragma unroll
        for (int j = 0; j < kDistGroup; ++j)
            if (j == lane && f[j] >= 0) lb = ball_member_lb(sqrt(d2[j]), frad[f[j]]);
        if (lane < kDistGroup && base + lane < n) lbs[base + lane] = lb;
    }
    wave_sync();
    double T;
    const float th = ordered_to_float(theta[q]);
    if (res) {
        T = th > 0.0f ? sqrt((double)th) * 1.02 + 1e-3 : 1e-3;
    } else {
        const double g = 2.0 * (double)(dim + 4) * 5.9604644775390625e-8;
        const double nq = (double)qnorm[q], nr = (double)__uint_as_float(*rmax);
        const double t2 = (double)th + nq * (1.0 + g + eq) + er * nr;
        T = t2 > 0.0 ? sqrt(t2) : 0.0;
    }
    // taken: rank among the passing candidates below cap
    int npass = 0;
    for (int i = 0; i < n; ++i) npass += (fs[i] >= 0 && (double)lbs[i] < T) ? 1 : 0;
    float lbmin = INFINITY;
    for (int i = lane; i < n; i += kWave) {
        const int f = fs[i];
        if (f < 0) continue;
        const float lb = lbs[i];
        bool take = false;
        if ((double)lb < T) {
            int rank = 0;
            for (int t = 0; t < n; ++t) {
                const int g2 = fs[t];
                const float l2 = lbs[t];
                rank += (g2 >= 0 && (double)l2 < T && (l2 < lb || (l2 == lb && g2 < f))) ? 1 : 0;
            }
            if (rank < cap) {
                out[o0 + 1 + rank] = f;
                take = true;
            }
        }
        if (!take) lbmin = fminf(lbmin, lb);
    }
    for (int i = min(npass, cap) + lane; i < Pb - 1; i += kWave) out[o0 + 1 + i] = -1;
    for (int o = 32; o > 0; o >>= 1) lbmin = fminf(lbmin, __shfl_xor(lbmin, o));
    if (lane == 0 && lbmin < lb_rest[q]) lb_rest[q] = lbmin;
}

// Every ball at once (one scan pass): for query q, the candidates are the
// non-empty balls of blocks chosen[q * P .. q * P + P), with FP64 centre
// distance d_f, radius rho_f, bound lb_f = d_f - rho_f and reach u_f = d_f +
// rho_f (every member of f is within u_f).  T_q = the smallest u_i such that
// the balls with u_j <= u_i hold at least kp members between them: those
// members are kp references within T_q, so the kp-th nearest distance is at
// most T_q and a ball with lb_f >= T_q holds no reference the scan would
// keep.  The balls with lb_f < T_q, nearest bound first (ties: lower ball),
// fill out[q * Pb ..] up to Pb entries (the rest -1); every other candidate
// lowers lb_rest[q] to lb_f.  (No such T_q, fewer than kp members in all:
// every candidate passes.)  At most kBallCand candidates.
constexpr int kSelBytes = 28;  // LDS bytes per candidate of ball_select_kernel
__global__ __launch_bounds__(64 * kBallWaves) void ball_select_kernel(
    const float *__restrict__ qpm, int pm_ld, int dim, int Q, int kp, const int *__restrict__ chosen, int P,
    const float *__restrict__ fcent, const double *__restrict__ frad, const int *__restrict__ fcnt,
    const int *__restrict__ fb_start, int *__restrict__ out, int Pb, float *__restrict__ lb_rest, int ncap) {
    // per wave: reach and radius (FP64), bound, ball and member count of
    // ncap candidates
    extern __shared__ __align__(16) unsigned char sel_smem[];
    __shared__ int blk_st[kBallWaves][kWave], blk_fl[kBallWaves][kWave];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int q = blockIdx.x * kBallWaves + wave;
    if (q >= Q) return;
    unsigned char *wbase = sel_smem + (size_t)wave * ncap * kSelBytes;
    double *us = reinterpret_cast<double *>(wbase);
    double *rh = reinterpret_cast<double *>(wbase + (size_t)ncap * 8);
    float *lbs = reinterpret_cast<float *>(wbase + (size_t)ncap * 16);
    int *fs = reinterpret_cast<int *>(wbase + (size_t)ncap * 20);
    int *ns = reinterpret_cast<int *>(wbase + (size_t)ncap * 24);
    const long long o0 = (long long)q * Pb;
    const float *qr = qpm + (long long)q * pm_ld;
    // candidates, block by block: the blocks' ball ranges read by one lane
    // each, placed by a prefix over the lanes; then one lane per candidate
    int n = 0;
    for (int j0 = 0; j0 < P; j0 += kWave) {
        const int j = j0 + lane;
        int fl = 0, cnt = 0;
        if (j < P) {
            const int b = chosen[(long long)q * P + j];
            if (b >= 0) {
                fl = fb_start[b];
                cnt = fb_start[b + 1] - fl;
            }
        }
        int pre = cnt;  // inclusive prefix over the lanes
        for (int o = 1; o < kWave; o <<= 1) {
            const int t = __shfl_up(pre, o);
            if (lane >= o) pre += t;
        }
        const int total = __shfl(pre, kWave - 1);
        const int nj = min(kWave, P - j0);
        // each block's first candidate slot and first ball, read below by
        // lanes of a divergent loop (so through LDS, not lane shuffles)
        if (lane < nj) {
            blk_st[wave][lane] = pre - cnt;
            blk_fl[wave][lane] = fl;
        }
        wave_sync();
        for (int i = lane; i < total; i += kWave) {
            // the block holding candidate i: the last one whose slots start at or before i
            int fb = 0, off = 0;
            for (int t = 0; t < nj; ++t) {
                const int st = blk_st[wave][t];
                if (st <= i) {
                    off = st;
                    fb = blk_fl[wave][t];
                }
            }
            const int f = fb + i - off;
            const double rho = frad[f];
            const bool live = rho >= 0.0;
            fs[n + i] = live ? f : -1;
            ns[n + i] = live ? fcnt[f] : 0;
            rh[n + i] = rho;
        }
        wave_sync();  // the next chunk rewrites blk_st / blk_fl
        n += total;
    }
    wave_sync();
    for (int base = 0; base < n; base += kDistGroup) {
        int f[kDistGroup];
#pragma unroll
        for (int j = 0; j < kDistGroup; ++j) f[j] = base + j < n ? fs[base + j] : -1;
        double d2[kDistGroup];
        wave_dist2_group(qr, fcent, pm_ld, dim, f, d2, lane);
        float lb = INFINITY;
        double u = INFINITY;
#pragma unroll
        for (int j = 0; j < kDistGroup; ++j)
            if (j == lane && f[j] >= 0) {
                const double d = sqrt(d2[j]), rho = rh[base + j];
                lb = ball_member_lb(d, rho);
                u = d * (1.0 + 1e-12) + rho;
            }
        if (lane < kDistGroup && base + lane < n) {
            lbs[base + lane] = lb;
            us[base + lane] = u;
        }
    }
    wave_sync();
    // T_q over the lanes' candidates, then the smallest
    double T = INFINITY;
    for (int i = lane; i < n; i += kWave) {
        if (fs[i] < 0) continue;
        const double ui = us[i];
        long long members = 0;
        for (int t = 0; t < n; ++t) members += us[t] <= ui ? ns[t] : 0;
        if (members >= kp) T = fmin(T, ui);
    }
    for (int o = 32; o > 0; o >>= 1) T = fmin(T, __shfl_xor(T, o));
    int npass = 0;
    for (int i = 0; i < n; ++i) npass += (fs[i] >= 0 && (double)lbs[i] < T) ? 1 : 0;
    float lbmin = INFINITY;
    for (int i = lane; i < n; i += kWave) {
        const int f = fs[i];
        if (f < 0) continue;
        const float lb = lbs[i];
        bool take = false;
        if ((double)lb < T) {
            int rank = 0;
            for (int t = 0; t < n; ++t) {
                const int g2 = fs[t];
                const float l2 = lbs[t];
                rank += (g2 >= 0 && (double)l2 < T && (l2 < lb || (l2 == lb && g2 < f))) ? 1 : 0;
            }
            if (rank < Pb) {
                out[o0 + rank] = f;
                take = true;
            }
        }
        if (!take) lbmin = fminf(lbmin, lb);
    }
    for (int i = min(npass, Pb) + lane; i < Pb; i += kWave) out[o0 + i] = -1;
    for (int o = 32; o > 0; o >>= 1) lbmin = fminf(lbmin, __shfl_xor(lbmin, o));
    if (lane == 0 && lbmin < lb_rest[q]) lb_rest[q] = lbmin;
}

// The same selection with one candidate per lane (at most 128 candidates,
// two per lane, and at most 64 probed blocks): lane i computes candidate i's
// FP64 centre distance over the dimensions in order, and the reach test,
// the pass count and the ranks run as wave-uniform loops over the
// candidates, reading each one's values from its lane (readlane), so no
// LDS round trips sit on the critical path.  Same candidates, bounds, T_q
// and ranks as ball_select_kernel (the FP64 sums differ only in their
// summation order, far inside the 1e-12 allowance).
// theta (nullable): each query's scan threshold is seeded from T_q.  The
// kp-th nearest distance is at most T_q, and the scan's key of a reference
// within T_q (||r||^2 - 2 r.q with error at most eq ||q||^2 + er max||r||^2)
// is below T_q^2 - ||q||^2 (1 - g) + eq ||q||^2 + er max||r||^2, widened and
// rounded up: the scan then starts from that threshold instead of +inf.
// Any starting threshold keeps the certificate sound (the final threshold
// is no larger than any the scan rejected against); a tight one saves the
// candidates of young lists.
constexpr int kSelLaneCand = 2;  // candidates per lane
__device__ __forceinline__ double readlane_f64(double x, int l) {
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
    return __hiloint2double(hi, lo);
}
__global__ __launch_bounds__(64 * kBallWaves) void ball_select_lanes_kernel(
    const float *__restrict__ qpm, int pm_ld, int dim, int Q, int kp, const int *__restrict__ chosen, int P,
    const float *__restrict__ fcent, const double *__restrict__ frad, const int *__restrict__ fcnt,
    const int *__restrict__ fb_start, int *__restrict__ out, int Pb, float *__restrict__ lb_rest,
    const float *__restrict__ qnorm, const unsigned *__restrict__ rmax, double eq, double er,
    unsigned *__restrict__ theta) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int q = __builtin_amdgcn_readfirstlane(blockIdx.x * kBallWaves + wave);
    if (q >= Q) return;  // (no workgroup barrier in this kernel)
    const long long o0 = (long long)q * Pb;
    // lane j < P: block j's first ball, its ball count and the candidates
    // before it (exclusive prefix)
    int fl = 0, cnt = 0;
    if (lane < P) {
        const int b = chosen[(long long)q * P + lane];
        if (b >= 0) {
            fl = fb_start[b];
            cnt = fb_start[b + 1] - fl;
        }
    }
    int pre = cnt;
    for (int o = 1; o < kWave; o <<= 1) {
        const int t = __shfl_up(pre, o);
        if (lane >= o) pre += t;
    }
    const int n = __builtin_amdgcn_readlane(pre, kWave - 1);
    const int excl = pre - cnt;
    // candidate i = lane + 64 k: its block is the last j with excl_j <= i
    int cf[kSelLaneCand], cm[kSelLaneCand];
    double crho[kSelLaneCand];
#pragma unroll
    for (int k = 0; k < kSelLaneCand; ++k) {
        const int i = lane + kWave * k;
        int first = 0, off = 0;
        for (int j = 0; j < P; ++j) {
            const int ej = __builtin_amdgcn_readlane(excl, j);
            const int fj = __builtin_amdgcn_readlane(fl, j);
            if (ej <= i) {
                off = ej;
                first = fj;
            }
        }
        cf[k] = -1;
        cm[k] = 0;
        crho[k] = -1.0;
        if (i < n) {
            const int f = first + i - off;
            const double rho = frad[f];
            if (rho >= 0.0) {
                cf[k] = f;
                cm[k] = fcnt[f];
                crho[k] = rho;
            }
        }
    }
    // FP64 centre distances, one candidate per lane
    const float *qr = qpm + (long long)q * pm_ld;
    double d2[kSelLaneCand];
#pragma unroll
    for (int k = 0; k < kSelLaneCand; ++k) d2[k] = 0.0;
    const bool two = n > kWave;  // wave-uniform
    for (int d = 0; d < dim; d += 4) {
        const float4 x = *reinterpret_cast<const float4 *>(qr + d);
        const double x0 = x.x, x1 = d + 1 < dim ? (double)x.y : 0.0, x2 = d + 2 < dim ? (double)x.z : 0.0,
                     x3 = d + 3 < dim ? (double)x.w : 0.0;
#pragma unroll
        for (int k = 0; k < kSelLaneCand; ++k) {
            if (k > 0 && !two) break;
            const int f = cf[k] >= 0 ? cf[k] : 0;
            const float4 c = *reinterpret_cast<const float4 *>(fcent + (long long)f * pm_ld + d);
            double t = x0 - (double)c.x;
            d2[k] = fma(t, t, d2[k]);
            t = d + 1 < dim ? x1 - (double)c.y : 0.0;
            d2[k] = fma(t, t, d2[k]);
            t = d + 2 < dim ? x2 - (double)c.z : 0.0;
            d2[k] = fma(t, t, d2[k]);
            t = d + 3 < dim ? x3 - (double)c.w : 0.0;
            d2[k] = fma(t, t, d2[k]);
        }
    }
    float clb[kSelLaneCand];
    double cu[kSelLaneCand];
#pragma unroll
    for (int k = 0; k < kSelLaneCand; ++k) {
        const double dd = sqrt(d2[k]);
        clb[k] = cf[k] >= 0 ? ball_member_lb(dd, crho[k]) : INFINITY;
        cu[k] = cf[k] >= 0 ? dd * (1.0 + 1e-12) + crho[k] : INFINITY;
    }
    // T_q: the smallest reach u_i whose balls u_t <= u_i hold >= kp members
    long long mem[kSelLaneCand] = {};
    for (int t = 0; t < n; ++t) {
        const int kt = t / kWave, lt = t - kt * kWave;  // wave-uniform
        const double ut = readlane_f64(kt ? cu[kSelLaneCand - 1] : cu[0], lt);
        const int mt = __builtin_amdgcn_readlane(kt ? cm[kSelLaneCand - 1] : cm[0], lt);
#pragma unroll
        for (int k = 0; k < kSelLaneCand; ++k) mem[k] += ut <= cu[k] ? mt : 0;
    }
    double T = INFINITY;
#pragma unroll
    for (int k = 0; k < kSelLaneCand; ++k)
        if (cf[k] >= 0 && mem[k] >= kp) T = fmin(T, cu[k]);
    for (int o = 32; o > 0; o >>= 1) T = fmin(T, __shfl_xor(T, o));
    if (theta && lane == 0) {  // (every query's threshold is written: +inf without a reach)
        unsigned tv = 0xFF800000u;
        if (T < INFINITY) {
            const double nq = (double)qnorm[q], nr = (double)__uint_as_float(*rmax);
            const double g = 2.0 * (double)(dim + 4) * 5.9604644775390625e-8;
            const double t =
                T * T * (1.0 + 1e-6) - nq * (1.0 - g - 1e-6) + (eq * nq + er * nr) * (1.0 + 1e-6) + 1e-30;
            float tf = (float)t;
            if ((double)tf <= t) tf = nextafterf(tf, INFINITY);
            tv = float_to_ordered(tf);
        }
        theta[q] = tv;
    }
    // passing candidates (bound below T_q), ranked by (bound, ball)
    bool pass[kSelLaneCand];
    int npass = 0;
#pragma unroll
    for (int k = 0; k < kSelLaneCand; ++k) {
        pass[k] = cf[k] >= 0 && (double)clb[k] < T;
        npass += __popcll(__ballot(pass[k]));
    }
    int rank[kSelLaneCand] = {};
    for (int t = 0; t < n; ++t) {
        const int kt = t / kWave, lt = t - kt * kWave;
        const int lbits = __builtin_amdgcn_readlane(__float_as_int(kt ? clb[kSelLaneCand - 1] : clb[0]), lt);
        const int ft = __builtin_amdgcn_readlane(kt ? cf[kSelLaneCand - 1] : cf[0], lt);
        const float l2 = __int_as_float(lbits);
        if (ft < 0 || !((double)l2 < T)) continue;  // wave-uniform
#pragma unroll
        for (int k = 0; k < kSelLaneCand; ++k) rank[k] += (l2 < clb[k] || (l2 == clb[k] && ft < cf[k])) ? 1 : 0;
    }
    float lbmin = INFINITY;
#pragma unroll
    for (int k = 0; k < kSelLaneCand; ++k) {
        if (cf[k] < 0) continue;
        if (pass[k] && rank[k] < Pb)
            out[o0 + rank[k]] = cf[k];
        else
            lbmin = fminf(lbmin, clb[k]);
    }
    for (int i = min(npass, Pb) + lane; i < Pb; i += kWave) out[o0 + i] = -1;
    for (int o = 32; o > 0; o >>= 1) lbmin = fminf(lbmin, __shfl_xor(lbmin, o));
    if (lane == 0 && lbmin < lb_rest[q]) lb_rest[q] = lbmin;
}

// Invert the (query -> blocks) lists, probes [p0, p1) of each query's P, into
// per-block query lists and scan tasks of up to 32 queries.  Step 1: count
// pairs per block.
__global__ void pair_count_kernel(const int *__restrict__ chosen, int Q, int P, int p0, int p1, int nu,
                                  int *__restrict__ counts) {
    const int w = p1 - p0;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < (long long)Q * w;
         e += (long long)gridDim.x * blockDim.x) {
        const long long q = e / w;
        const int b = chosen[q * P + p0 + (e - q * w)];
        if (b >= 0 && b < nu) atomicAdd(&counts[b], 1);
    }
}

// Step 2: place each pair in its block's list (qlist, grouped by block with
// each block's group starting at a multiple of 32) and record the heap row
// it will produce: the pair's slot in the list, plus row_base, is its
// output row.
__global__ void pair_scatter_kernel(const int *__restrict__ chosen, int Q, int P, int p0, int p1, int nu,
                                    int *__restrict__ cursor, int *__restrict__ qlist,
                                    int *__restrict__ rows, int row_base) {
    const int w = p1 - p0;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < (long long)Q * w;
         e += (long long)gridDim.x * blockDim.x) {
        const long long q = e / w;
        const long long pi = q * P + p0 + (e - q * w);
        const int b = chosen[pi];
        if (b < 0 || b >= nu) {  // (no unit: units outside [0, nu) are never read)
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

// Steps 1b and 3 in one launch: the units' list starts, the cursors the
// scatter advances, and the scan tasks.  Unit b's pairs take
// ceil(count[b] / qpt) tasks (qpt queries a task: 32, or 16 for merged-list
// tasks), so its list starts at qpt x the exclusive prefix of the task
// counts (slot == output row) and the tasks of all units fill [0, T) with
// no gaps; slots [T, max_tasks) are written empty (they scan nothing).
// Units are blocks (tile-aligned starts) or balls (any start: a task's tiles
// begin at the start rounded down to 32).  Every workgroup scans all nu
// counts (a few KB, from L2) and writes the units and tail slots of its own
// share.
constexpr int kUnitTaskThreads = 1024, kUnitTaskIpt = 8, kUnitTaskWgs = 16;
__global__ __launch_bounds__(kUnitTaskThreads) void unit_tasks_kernel(
    const int *__restrict__ counts, int nu, int qpt, const int *__restrict__ unit_start,
    const int *__restrict__ unit_count, int *__restrict__ qstart, int *__restrict__ cursor,
    ScanTask *__restrict__ tasks, int max_tasks, int *__restrict__ clear) {
    constexpr int nw = kUnitTaskThreads / kWave;
    __shared__ int wsum[nw];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const int share = (nu + gridDim.x - 1) / gridDim.x;
    const int own0 = min(nu, (int)blockIdx.x * share), own1 = min(nu, own0 + share);
    int carry = 0;  // tasks of the units before the chunk
    for (int c0 = 0; c0 < nu; c0 += kUnitTaskThreads * kUnitTaskIpt) {
        const int i0 = c0 + tid * kUnitTaskIpt;
        int cnt[kUnitTaskIpt], tot = 0;
#pragma unroll
        for (int i = 0; i < kUnitTaskIpt; ++i) {
            cnt[i] = i0 + i < nu ? counts[i0 + i] : 0;
            tot += (cnt[i] + qpt - 1) / qpt;
        }
        // exclusive prefix of the threads' task counts: within the wave,
        // then the earlier waves' totals
        int v = tot;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const int x = __shfl_up(v, off);
            if (lane >= off) v += x;
        }
        if (lane == kWave - 1) wsum[wave] = v;
        __syncthreads();
        int before = 0, chunk = 0;
        for (int w = 0; w < nw; ++w) {
            const int x = wsum[w];
            before += w < wave ? x : 0;
            chunk += x;
        }
        __syncthreads();  // wsum is written again by the next chunk
        int t = carry + before + v - tot;  // the first task of the thread's units
#pragma unroll
        for (int i = 0; i < kUnitTaskIpt; ++i) {
            const int b = i0 + i;
            const int nt = (cnt[i] + qpt - 1) / qpt;
            if (b >= own0 && b < own1) {
                qstart[b] = t * qpt;
                cursor[b] = t * qpt;
                const int lo = unit_start[b], hi = lo + unit_count[b];
                for (int j = 0; j < nt && t + j < max_tasks; ++j) {
                    ScanTask tk;
                    tk.ref_lo = lo;
                    tk.ref_begin = lo / kRefsPerTile * kRefsPerTile;
                    tk.ref_end = hi;
                    tk.qoff = (t + j) * qpt;
                    tk.qcount = min(qpt, cnt[i] - j * qpt);
                    tk.unit = b;
                    tasks[t + j] = tk;
                }
            }
            t += nt;
        }
        carry += chunk;
    }
    if (blockIdx.x == 0 && tid == 0) {
        qstart[nu] = carry * qpt;
        if (clear) *clear = 0;
    }
    for (int t = carry + (int)blockIdx.x * kUnitTaskThreads + tid; t < max_tasks;
         t += (int)gridDim.x * kUnitTaskThreads) {
        ScanTask tk;
        tk.ref_begin = 0;
        tk.ref_end = 0;
        tk.qoff = 0;
        tk.qcount = 0;
        tk.ref_lo = 0;
        tk.unit = 0;
        tasks[t] = tk;
    }
}

// Block of every reference point (original order): one workgroup per block
// walks the block's positions.  Padding positions (perm < 0) are skipped.
__global__ void ref_block_kernel(const int *__restrict__ block_start,
                                 const int *__restrict__ block_count, const int *__restrict__ perm,
                                 int *__restrict__ out) {
    const int b = blockIdx.x;
    const int s0 = block_start[b], n = block_count[b];
    for (int p = threadIdx.x; p < n; p += blockDim.x) {
        const int r = perm[s0 + p];
        if (r >= 0) out[r] = b;
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

// Batched search: position i of a concatenation of point sets (set p at
// starts[p] .. starts[p] + counts[p], each start a multiple of 32) maps to
// the set's own index i - starts[p]; positions between sets get -1.  One
// thread per position, the set found by binary search over starts.
__global__ void batch_perm_kernel(const int *__restrict__ starts, const int *__restrict__ counts,
                                  int nsets, int n_pos, int *__restrict__ perm) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_pos; i += gridDim.x * blockDim.x) {
        int lo = 0, hi = nsets - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (starts[mid] <= i)
                lo = mid;
            else
                hi = mid - 1;
        }
        const int local = i - starts[lo];
        perm[i] = local < counts[lo] ? local : -1;
    }
}

// ---------------------------------------------------------------------------
// k-means++ seeding (D^2 sampling, Arthur & Vassilvitskii, SODA 2007)
// ---------------------------------------------------------------------------

// splitmix64 finaliser of x, as a double uniform in [0, 1).
__device__ inline double hash_unit(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (double)(x >> 11) * 0x1.0p-53;
}

// Squared distance of two point-major rows over `dim` coordinates (FP32;
// rows are 16-byte aligned, pad coordinates are not read).
__device__ inline float row_sqdist(const float *__restrict__ a, const float *__restrict__ b, int dim) {
    float s = 0.0f;
    int d = 0;
    for (; d + 4 <= dim; d += 4) {
        const float4 x = *reinterpret_cast<const float4 *>(a + d);
        const float4 y = *reinterpret_cast<const float4 *>(b + d);
        const float t0 = x.x - y.x, t1 = x.y - y.y, t2 = x.z - y.z, t3 = x.w - y.w;
        s = fmaf(t0, t0, s);
        s = fmaf(t1, t1, s);
        s = fmaf(t2, t2, s);
        s = fmaf(t3, t3, s);
    }
    for (; d < dim; ++d) {
        const float t = a[d] - b[d];
        s = fmaf(t, t, s);
    }
    return s;
}

// One wave: weighted pick over v[0..n) (v >= 0).  Lane l sums the
// contiguous range [l*per, (l+1)*per); the pick is the first index whose
// running sum exceeds the target (target = u * total if `relative`, else
// u itself), or the last positive entry if rounding leaves none; -1 when
// every entry is 0.  *before receives the running sum ahead of the pick.
// *target_out (if given) receives the target.  All lanes return the same
// values.
template <class T>
__device__ int wave_pick(const T *__restrict__ v, int n, double u, bool relative, double *before,
                         double *target_out) {
    const int lane = threadIdx.x & 63;
    const int per = (n + 63) / 64;
    const int i0 = min(n, lane * per), i1 = min(n, i0 + per);
    double s = 0.0;
    int lastpos = -1;
    for (int i = i0; i < i1; ++i) {
        const double x = (double)v[i];
        s += x;
        if (x > 0.0) lastpos = i;
    }
    double incl = s;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
    }
    double excl = __shfl_up(incl, 1);
    if (lane == 0) excl = 0.0;
    const double total = __shfl(incl, 63);
    const double target = relative ? u * total : u;
    if (target_out) *target_out = target;
    const unsigned long long cross = __ballot(i1 > i0 && incl > target);
    int pick = -1;
    double pb = 0.0;
    if (cross) {
        const int L = __ffsll((long long)cross) - 1;
        int p = -1;
        double b = 0.0;
        if (lane == L) {
            double run = excl;
            int lp = -1;
            double lpb = 0.0;
            for (int i = i0; i < i1; ++i) {
                const double x = (double)v[i];
                if (x > 0.0) {
                    lp = i;
                    lpb = run;
                }
                if (run + x > target && x > 0.0) {
                    p = i;
                    b = run;
                    break;
                }
                run += x;
            }
            if (p < 0) {
                p = lp;
                b = lpb;
            }
        }
        pick = __shfl(p, L);
        pb = __shfl(b, L);
    }
    if (pick < 0) {
        const unsigned long long has = __ballot(lastpos >= 0);
        if (has) {
            const int L = 63 - __clzll((long long)has);
            pick = __shfl(lastpos, L);
            pb = __shfl(excl, L);
        }
    }
    if (before) *before = pb;
    return pick;
}

// Sample i of the seeding set is point floor(i * n / m) (an even stride).
__device__ inline long long sample_point(int i, long long n, int m) { return (long long)i * n / m; }

__global__ void kpp_init_kernel(float *__restrict__ mind, int m, int *__restrict__ cidx,
                                unsigned *__restrict__ ticket, unsigned long long seed) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x)
        mind[i] = INFINITY;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        cidx[0] = min(m - 1, (int)(hash_unit(seed) * m));
        *ticket = 0u;
    }
}

// Step t: fold centre t-1 into every sample's squared distance to its
// nearest centre, then (in the workgroup that finishes last) draw centre t
// with probability proportional to that distance.  Each workgroup owns a
// contiguous sample range and publishes its sum; the draw picks a range by
// those sums and a sample inside it.
__global__ __launch_bounds__(256) void kpp_step_kernel(const float *__restrict__ pm, int ld, int dim,
                                                       long long n, int m, float *__restrict__ mind,
                                                       int *__restrict__ cidx, int t,
                                                       unsigned long long seed,
                                                       double *__restrict__ wsum,
                                                       unsigned *__restrict__ ticket) {
    extern __shared__ float csh[];  // the new centre, ld floats
    __shared__ double red[256];
    __shared__ int is_last;
    const int tid = threadIdx.x;
    const float *c = pm + sample_point(cidx[t - 1], n, m) * ld;
    for (int d = tid; d < ld; d += blockDim.x) csh[d] = c[d];
    __syncthreads();
    const int per = (m + gridDim.x - 1) / gridDim.x;
    const int i0 = min(m, (int)blockIdx.x * per), i1 = min(m, i0 + per);
    double acc = 0.0;
    for (int i = i0 + tid; i < i1; i += blockDim.x) {
        const float d2 = row_sqdist(pm + sample_point(i, n, m) * ld, csh, dim);
        const float v = fminf(mind[i], d2);
        mind[i] = v;
        acc += (double)v;
    }
    red[tid] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    if (tid == 0) wsum[blockIdx.x] = red[0];
    __threadfence();
    __syncthreads();
    if (tid == 0) is_last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    if (tid < 64) {
        const double u = hash_unit(seed ^ ((unsigned long long)t * 0xD1B54A32D192ED03ull));
        double before = 0.0, target = 0.0;
        const int w = wave_pick(wsum, gridDim.x, u, true, &before, &target);
        int pick = t % m;  // every sample already a centre: any choice
        if (w >= 0) {
            const int j0 = min(m, w * per), j1 = min(m, j0 + per);
            const int r = wave_pick(mind + j0, j1 - j0, target - before, false, nullptr, nullptr);
            pick = r >= 0 ? j0 + r : j0;
        }
        if (tid == 0) {
            cidx[t] = pick;
            *ticket = 0u;
        }
    }
}

// Centres from the drawn samples: cent[b] = row of sample cidx[b].
__global__ void kpp_gather_kernel(const float *__restrict__ pm, int ld, long long n, int m,
                                  const int *__restrict__ cidx, int nb, float *__restrict__ cent) {
    const long long total = (long long)nb * ld;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long b = e / ld, d = e - b * ld;
        cent[e] = pm[sample_point(cidx[b], n, m) * ld + d];
    }
}

// ---------------------------------------------------------------------------
// Blocks as balanced groups of fine balls
// ---------------------------------------------------------------------------

// Fine ball f (in index order) joins the nearest of its candidate blocks
// cand[j * nf + f], j < kc (ascending distance, as the engine returns them)
// whose load stays within `cap` points with it; when none has room, the
// least-loaded candidate (ties: nearer).  Empty balls join their nearest
// candidate.  One wave, sequential over the balls; loads live in LDS.
__global__ __launch_bounds__(64) void group_fine_kernel(const int *__restrict__ cand, int kc, int nf,
                                                        const int *__restrict__ fcount, int nb,
                                                        int cap, int *__restrict__ fblock) {
    extern __shared__ int load[];
    const int lane = threadIdx.x;
    for (int b = lane; b < nb; b += 64) load[b] = 0;
    __syncthreads();
    for (int f = 0; f < nf; ++f) {
        const int c = fcount[f];
        const int b = lane < kc ? cand[(long long)lane * nf + f] : -1;
        const int l = b >= 0 ? load[b] : 0x7FFFFFFF;
        const unsigned long long fits = __ballot(b >= 0 && (c == 0 || l <= cap - c));
        int pick;
        if (fits) {
            pick = __shfl(b, __ffsll((long long)fits) - 1);
        } else {
            int bl = l, bj = lane;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                const int ol = __shfl_xor(bl, off), oj = __shfl_xor(bj, off);
                if (ol < bl || (ol == bl && oj < bj)) {
                    bl = ol;
                    bj = oj;
                }
            }
            pick = __shfl(b, bj);
        }
        if (lane == 0) {
            fblock[f] = pick;
            load[pick] += c;
        }
        __syncthreads();
    }
}

// Per block: point count and fine-ball count.
__global__ void block_count_kernel(const int *__restrict__ fblock, const int *__restrict__ fcount,
                                   int nf, int *__restrict__ bcount, int *__restrict__ bfine) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
        atomicAdd(&bcount[fblock[f]], fcount[f]);
        atomicAdd(&bfine[fblock[f]], 1);
    }
}

// Each fine ball takes a contiguous range of positions inside its block
// (fstart) and a slot in the block-sorted ball order (fine_sorted).
__global__ void fine_start_kernel(const int *__restrict__ fblock, const int *__restrict__ fcount,
                                  int nf, int *__restrict__ pos_cursor, int *__restrict__ fstart,
                                  int *__restrict__ slot_cursor, int *__restrict__ fine_sorted) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
        fstart[f] = atomicAdd(&pos_cursor[fblock[f]], fcount[f]);
        fine_sorted[atomicAdd(&slot_cursor[fblock[f]], 1)] = f;
    }
}

// Ball s of the block-sorted order (fine ball f = fine_sorted[s]): its
// centre row, block, and radius = max member distance (double, widened by
// a relative 1e-12 so that it bounds the exact distance; -1 when empty).
// One workgroup per ball; members are positions fstart[f] .. + fcount[f].
__global__ __launch_bounds__(256) void finalize_fine_kernel(const float *__restrict__ pm, int ld,
                                                            int dim, const float *__restrict__ fcent,
                                                            const int *__restrict__ fine_sorted,
                                                            const int *__restrict__ fstart,
                                                            const int *__restrict__ fcount,
                                                            const int *__restrict__ fblock,
                                                            float *__restrict__ cent_s,
                                                            double *__restrict__ rad_s,
                                                            int *__restrict__ block_s) {
    __shared__ double red[256];
    const int sidx = blockIdx.x, tid = threadIdx.x;
    const int f = fine_sorted[sidx];
    const int n = fcount[f];
    const float *c = fcent + (long long)f * ld;
    for (int d = tid; d < ld; d += blockDim.x) cent_s[(long long)sidx * ld + d] = d < dim ? c[d] : 0.0f;
    double best = 0.0;
    const long long s0 = fstart[f];
    for (int j = tid; j < n; j += blockDim.x) {
        const float *r = pm + (s0 + j) * ld;
        double acc = 0.0;
        for (int d = 0; d < dim; ++d) {
            const double t = (double)r[d] - (double)c[d];
            acc += t * t;
        }
        best = fmax(best, acc);
    }
    red[tid] = best;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] = fmax(red[tid], red[tid + w]);
        __syncthreads();
    }
    if (tid == 0) {
        rad_s[sidx] = n > 0 ? sqrt(red[0]) * (1.0 + 1e-12) : -1.0;
        block_s[sidx] = fblock[f];
    }
}

// Radii of the balls of a loaded index, refitted to the reference set it is
// loaded over: every point of block b joins the ball of b whose centre is
// nearest (FP64), and a ball's radius becomes the largest such member
// distance.  Every point then lies within the radius of a ball of its own
// block, which is all the block lower bounds need, whatever the file held.
// One workgroup per block; d2bits: nf zeroed words (FP64 bit patterns of
// non-negative squared distances order like the values); member: nf zeroed
// flags.
__global__ __launch_bounds__(256) void ball_refit_kernel(const float *__restrict__ pm, int ld, int dim,
                                                         const int *__restrict__ block_start,
                                                         const int *__restrict__ block_count,
                                                         const int *__restrict__ fb_start,
                                                         const float *__restrict__ fcent,
                                                         unsigned long long *d2bits, int *member) {
    const int b = blockIdx.x;
    const long long p0 = block_start[b];
    const int n = block_count[b], f0 = fb_start[b], f1 = fb_start[b + 1];
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const float *r = pm + (p0 + j) * ld;
        double best = INFINITY;
        int bf = -1;
        for (int f = f0; f < f1; ++f) {
            const float *c = fcent + (long long)f * ld;
            double acc = 0.0;
            for (int d = 0; d < dim; ++d) {
                const double t = (double)r[d] - (double)c[d];
                acc += t * t;
            }
            if (acc < best) {
                best = acc;
                bf = f;
            }
        }
        if (bf >= 0) {
            atomicMax(d2bits + bf, (unsigned long long)__double_as_longlong(best));
            atomicOr(member + bf, 1);
        }
    }
}

// Radii of balls with known position ranges: one workgroup per ball, the
// largest FP64 distance from its centre to a position of its range.
__global__ __launch_bounds__(256) void ball_range_radius_kernel(const float *__restrict__ pm, int ld, int dim,
                                                                const float *__restrict__ fcent,
                                                                const int *__restrict__ ball_pos,
                                                                const int *__restrict__ ball_cnt, double *rad) {
    __shared__ double red[256];
    const int f = blockIdx.x, tid = threadIdx.x;
    const long long p0 = ball_pos[f];
    const int n = ball_cnt[f];
    const float *c = fcent + (long long)f * ld;
    double best = 0.0;
    for (int j = tid; j < n; j += blockDim.x) {
        const float *r = pm + (p0 + j) * ld;
        double acc = 0.0;
        for (int d = 0; d < dim; ++d) {
            const double t = (double)r[d] - (double)c[d];
            acc += t * t;
        }
        best = fmax(best, acc);
    }
    red[tid] = best;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] = fmax(red[tid], red[tid + w]);
        __syncthreads();
    }
    if (tid == 0) rad[f] = n > 0 ? sqrt(red[0]) * (1.0 + 1e-12) : -1.0;
}

__global__ void ball_refit_finish_kernel(const unsigned long long *__restrict__ d2bits,
                                         const int *__restrict__ member, int nf, double *rad) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x)
        rad[f] = member[f] ? sqrt(__longlong_as_double((long long)d2bits[f])) * (1.0 + 1e-12) : -1.0;
}

// ---------------------------------------------------------------------------
// Search: block lower bounds
// ---------------------------------------------------------------------------

// Lower bound of every reference of block b for query q (distance, not
// squared, rounded down to float): the smallest over the block's fine balls
// f of max(0, |q - c_f| - rho_f), empty balls excluded; lb[b][q] (block-
// major, query_nb per row) must hold +inf on entry and keeps +inf for a
// block without balls.  |q - c_f|^2 comes from ||q||^2 + ||c||^2 - 2 q.c on
// FP32 MFMA, lowered by the error bound 1.5 g (||q||^2 + ||c||^2 + 2|q||c|),
// g = 2 (dim + 4) 2^-24.  Balls are in block-sorted order (block_s
// ascending, block b's balls at fb_start[b] .. fb_start[b+1]).  Each wave
// computes a 32 (balls) x 32 (queries) tile (k-step kk reads coordinate
// kk + h*KH on lane half h, KH = half the padded dimension), stages it in
// LDS, and each lane then folds 16 consecutive balls of one query into
// per-block minima: a block wholly inside those 16 balls is stored, one
// that extends past them is merged with an atomic min on the float bits
// (non-negative floats order like their bit patterns).
// max(0, |q - c| - rho) rounded down, in FP32 arithmetic.  nq, nc, dot
// carry the FP32 norm and MFMA errors, within 1.5 g S for
// S = (sqrt(nq) + sqrt(nc))^2 (g15 >= 1.5 g; the errors themselves are
// within g S, which leaves 0.5 g S spare).  The allowance for them is
// c S_f with c = (g15 + 2^-21)(1 + 2^-20) and S_f = (sq + sc)^2 from square
// roots rounded up; S_f <= 2 sq^2 + 2 sc^2, so c (2 sq^2 + 2 sc^2) covers it,
// split between the two ends: each end's share is taken off its own squared
// norm once (ball_norm_share: pq = nq - 2 c sq^2 for the query, pc for the
// ball, rounded down), and a pair's d2 = pq + pc - 2 dot costs two roundings,
// each within 1.01 u S (u = 2^-24; S bounds nq + nc, 2|dot| and the partial
// results), covered by the 2^-21 S in the allowance and the spare.  The
// hardware square root (within 1 ulp) is lowered by 2^-21 relative, the
// radius comes rounded up (ru < 0: empty ball), and the difference is
// lowered by 2^-22 relative.
__device__ __forceinline__ float ball_norm_share(float n, float s, float g15) {
    const double c = ((double)g15 + 0x1p-21) * (1.0 + 0x1p-20);
    const double v = (double)n - 2.0 * c * (double)s * (double)s;
    const double w = v - fabs(v) * 0x1p-50;  // (below the exact value despite the FP64 roundings)
    float f = (float)w;
    if ((double)f > w) f = nextafterf(f, -INFINITY);
    return f;
}
// (the bound is final(raw): final is monotone, so the minimum over balls of
// the bounds is final of the minimum of the raw values, bit for bit)
__device__ __forceinline__ float ball_bound_raw(float pq, float pc, float dot, float ru) {
    const float d2 = fmaxf(fmaf(-2.0f, dot, pq + pc), 0.0f);
    return __builtin_amdgcn_sqrtf(d2) * (1.0f - 0x1p-21f) - ru;
}
__device__ __forceinline__ float ball_bound_final(float v) {
    return fmaxf(v, 0.0f) * (1.0f - 0x1p-22f);  // (ru = -inf: +inf)
}
__device__ __forceinline__ float ball_lower_bound_core(float pq, float pc, float dot, float ru) {
    return ball_bound_final(ball_bound_raw(pq, pc, dot, ru));
}
// (the per-ball forms: the ball's share computed on the spot)
__device__ __forceinline__ float ball_lower_bound(float pq, float nc, float sc, float dot, float ru, float g15) {
    if (ru < 0.0f) return INFINITY;
    return ball_lower_bound_core(pq, ball_norm_share(nc, sc, g15), dot, ru);
}

// Per ball: the square root of the centre's squared norm and the radius,
// both rounded up to float (radius -1 for an empty ball), for the bounds.
__global__ void ball_aux_kernel(const float *__restrict__ fnorm, const double *__restrict__ frad,
                                int nf, float *__restrict__ fsq, float *__restrict__ fradf) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
        fsq[f] = nextafterf(sqrtf(fnorm[f]), INFINITY);
        const double rad = frad[f];
        float ru = (float)rad;
        if ((double)ru < rad) ru = nextafterf(ru, INFINITY);
        fradf[f] = rad >= 0.0 ? ru : -1.0f;
    }
}

// 1.5 g, g = 2 (dim + 4) 2^-24, rounded up to float.
__device__ __forceinline__ float ball_error_coeff(int dim) {
    return (float)(3.0 * (double)(dim + 4) * 5.9604644775390625e-8) * (1.0f + 0x1p-20f);
}

typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));

// Ball centres as bfloat16 hi/lo pairs for the three-term bound products:
// ball f, 16-dim chunk c at out[(f*nch + c)*32]: 16 hi values (dims 16c ..
// 16c + 15, zero past dim), then the 16 lo values, lo = bf16(x - hi).
__global__ void ball_bf3_pack_kernel(const float *__restrict__ fcent, int pm_ld, int nf, int dim, int nch,
                                     unsigned short *__restrict__ out) {
    const long long n = (long long)nf * nch * 16;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(e % 16);
        const long long fc = e / 16;
        const int c = (int)(fc % nch), f = (int)(fc / nch);
        const int d = 16 * c + j;
        const float x = d < dim ? fcent[(long long)f * pm_ld + d] : 0.0f;
        const __bf16 hi = (__bf16)x;
        const __bf16 lo = (__bf16)(x - (float)hi);
        out[fc * 32 + j] = __builtin_bit_cast(unsigned short, hi);
        out[fc * 32 + 16 + j] = __builtin_bit_cast(unsigned short, lo);
    }
}

__global__ __launch_bounds__(256) void ball_lb_kernel(const float *__restrict__ qpm, int pm_ld, int Q,
                                                      int dim, const float *__restrict__ qnorm,
                                                      const float *__restrict__ fcent, int nf,
                                                      const float *__restrict__ fnorm,
                                                      const float *__restrict__ fsq,
                                                      const float *__restrict__ fradf,
                                                      const int *__restrict__ block_s,
                                                      const int *__restrict__ fb_start,
                                                      float *__restrict__ lb, int nb) {
    __shared__ float patch[4][32][33];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 31, h = lane >> 5;
    const int ntf = (nf + 31) / 32;
    const long long tile = (long long)blockIdx.x * 4 + wave;
    const int tq = (int)(tile / ntf), tf = (int)(tile - (long long)tq * ntf);
    const int q0 = tq * 32, f0 = tf * 32;
    const bool active = q0 < Q;
    if (active) {
        const int qa = min(q0 + col, Q - 1);  // B operand: query col
        const int fa = min(f0 + col, nf - 1); // A operand: ball col
        const float *qr = qpm + (long long)qa * pm_ld;
        const float *cr = fcent + (long long)fa * pm_ld;
        const int KH = (dim + 1) / 2;
        f32x16_t acc = {};
        if ((KH & 3) == 0 && (pm_ld & 3) == 0) {
            const float *ca = cr + h * KH, *qb = qr + h * KH;
            for (int k0 = 0; k0 < KH; k0 += 16) {
                float4 av[4], bv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const bool ok = k0 + 4 * u < KH;
                    av[u] = ok ? *reinterpret_cast<const float4 *>(ca + k0 + 4 * u) : make_float4(0, 0, 0, 0);
                    bv[u] = ok ? *reinterpret_cast<const float4 *>(qb + k0 + 4 * u) : make_float4(0, 0, 0, 0);
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
        // acc[i] is ball f0 + 8(i/4) + 4h + i%4 for query q0 + col
        const float g15 = ball_error_coeff(dim);
        const float nq = qnorm[qa];
        const float pq = ball_norm_share(nq, nextafterf(sqrtf(nq), INFINITY), g15);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int r = ((i >> 2) << 3) + (h << 2) + (i & 3);
            const int f = f0 + r;
            patch[wave][r][col] = f < nf ? ball_lower_bound(pq, fnorm[f], fsq[f], acc[i], fradf[f], g15)
                                         : INFINITY;
        }
    }
    __syncthreads();
    const int q = q0 + col;
    if (!active || q >= Q) return;
    const int lo = f0 + 16 * h, hi = min(nf, lo + 16);
    int cur = -1;
    float m = INFINITY;
    auto flush = [&](int b, float v) {
        if (b < 0) return;
        float *dst = lb + lb_at(b, q, nb);
        if (fb_start[b] >= lo && fb_start[b + 1] <= hi)
            *dst = v;
        else
            atomicMin(reinterpret_cast<unsigned *>(dst), __float_as_uint(v));
    };
    for (int f = lo; f < hi; ++f) {
        const int b = block_s[f];
        const float v = patch[wave][f - f0][col];
        if (b != cur) {
            flush(cur, m);
            cur = b;
            m = v;
        } else {
            m = fminf(m, v);
        }
    }
    flush(cur, m);
}

// The same bounds with the query operands in registers: a workgroup takes
// 128 queries (32 per wave) and `tpw` consecutive ball tiles [F0, F1)
// (fewer for few queries, so that the grid still covers the chip); every
// wave walks those tiles (so their loads are shared through L1) against its
// own 32 queries, whose half-rows (KH floats per lane, KH % 4 == 0,
// KH <= KHM) stay in VGPRs.  The workgroup first stages its tiles' per-ball
// data in LDS (squared norm, its root and radius rounded up, the block and
// its ball range).  The per-block fold runs in registers on lanes 0..31 (one
// query each): a tile's 32 bounds of a query sit on its two lane halves
// (rows 8(i/4) + 4h + i%4), the partner half's come over by a lane swap,
// and each row's block and that block's ball range come from the lane of the
// row by readlane.  The fold carries the open block from tile to tile, so a
// block is merged with an atomic min only when it extends past [F0, F1);
// every other block is stored once.
constexpr int kBallTilesPerWg = 8;  // tiles per workgroup for many queries
constexpr int kBallMinWgs = 1024;    // workgroups aimed for when there are few
//
// NCH > 0: the products are bfloat16 three-term (q_hi c_hi + q_hi c_lo +
// q_lo c_hi per 16-dim chunk, v_mfma_f32_32x32x16_bf16, FP32 accumulation)
// on the packed centres cbf; the accumulator layout is the FP32 MFMA's, so
// the bounds and the fold are shared.  g15_bf then replaces the FP32 error
// coefficient (the host's ball_error_coeff_bf3).

// Per-ball data of a workgroup's tiles, staged in LDS (rows past the last
// ball: radius -1, block -1).
struct BallStage {
    float nrm[kBallTilesPerWg * 32], sq[kBallTilesPerWg * 32], rad[kBallTilesPerWg * 32];
    int blk[kBallTilesPerWg * 32], bs[kBallTilesPerWg * 32], be[kBallTilesPerWg * 32];
};
__device__ __forceinline__ void ball_stage(BallStage &st, int F0, int F1, int tpw, const float *__restrict__ fnorm,
                                           const float *__restrict__ fsq, const float *__restrict__ fradf,
                                           const int *__restrict__ block_s, const int *__restrict__ fb_start) {
    for (int j = threadIdx.x; j < tpw * 32; j += blockDim.x) {
        const int f = F0 + j;
        const bool in = f < F1;
        const int b = in ? block_s[f] : -1;
        st.nrm[j] = in ? fnorm[f] : 0.0f;
        st.sq[j] = in ? fsq[f] : 0.0f;
        st.rad[j] = in ? fradf[f] : -1.0f;
        st.blk[j] = b;
        st.bs[j] = in ? fb_start[b] : 0;
        st.be[j] = in ? fb_start[b + 1] : 0;
    }
}

// A wave's query operands: lane (col, h) holds query qa's half-row (dims
// h*KH .. h*KH + KH - 1, FP32), or its hi / lo parts of dims 16c + 8h .. + 7
// per chunk c (bfloat16).
template <int KHM, int NCH>
__device__ __forceinline__ void ball_query_ops(const float *__restrict__ qpm, int pm_ld, int qa, int dim, int h,
                                               float4 (&bq)[NCH ? 1 : KHM / 4], bf16x8_t (&bh)[NCH ? NCH : 1],
                                               bf16x8_t (&bl)[NCH ? NCH : 1]) {
    const int KH = (dim + 1) / 2;
    if constexpr (NCH > 0) {
        const float *qr = qpm + (long long)qa * pm_ld;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const int d0 = 16 * c + 8 * h;
            float x[8];
            if (d0 + 8 <= dim && pm_ld % 4 == 0) {
                // eight dimensions inside the row: two 16-byte loads
                const float4 u = *reinterpret_cast<const float4 *>(qr + d0);
                const float4 w = *reinterpret_cast<const float4 *>(qr + d0 + 4);
                x[0] = u.x, x[1] = u.y, x[2] = u.z, x[3] = u.w;
                x[4] = w.x, x[5] = w.y, x[6] = w.z, x[7] = w.w;
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) x[e] = d0 + e < dim ? qr[d0 + e] : 0.0f;
            }
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const __bf16 hi = (__bf16)x[e];
                bh[c][e] = hi;
                bl[c][e] = (__bf16)(x[e] - (float)hi);
            }
        }
    } else {
        const float *qr = qpm + (long long)qa * pm_ld + h * KH;
#pragma unroll
        for (int u = 0; u < KHM / 4; ++u)
            bq[u] = 4 * u < KH ? *reinterpret_cast<const float4 *>(qr + 4 * u) : make_float4(0, 0, 0, 0);
    }
}

// Bounds of one 32-ball tile (balls f0 .. f0 + 31, whose squared norms, their
// roots and radii sit in LDS from tnrm / tsq / trad on) against a wave's 32
// queries: out[i] is ball f0 + 8(i/4) + 4h + i%4 for query col (+inf past
// the last ball).
template <int KHM, int NCH>
__device__ __forceinline__ void ball_tile_bounds(const float *tnrm, const float *tsq, const float *trad, int f0,
                                                 int nf, int pm_ld, int dim,
                                                 int col, int h, const float *__restrict__ fcent,
                                                 const unsigned short *__restrict__ cbf,
                                                 const float4 (&bq)[NCH ? 1 : KHM / 4],
                                                 const bf16x8_t (&bh)[NCH ? NCH : 1],
                                                 const bf16x8_t (&bl)[NCH ? NCH : 1], float pq, float g15,
                                                 float (&out)[16]) {
    const int KH = (dim + 1) / 2;
    const int fa = min(f0 + col, nf - 1);
    f32x16_t acc = {};
    if constexpr (NCH > 0) {
        const unsigned short *ca = cbf + (long long)fa * NCH * 32 + 8 * h;
        u32x4_t ca_hi[NCH], ca_lo[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            ca_hi[c] = *reinterpret_cast<const u32x4_t *>(ca + 32 * c);
            ca_lo[c] = *reinterpret_cast<const u32x4_t *>(ca + 32 * c + 16);
        }
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const bf16x8_t ah = __builtin_bit_cast(bf16x8_t, ca_hi[c]);
            const bf16x8_t al = __builtin_bit_cast(bf16x8_t, ca_lo[c]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh[c], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl[c], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh[c], acc, 0, 0, 0);
        }
    } else {
        const float *ca = fcent + (long long)fa * pm_ld + h * KH;
#pragma unroll
        for (int u = 0; u < KHM / 4; ++u) {
            if (4 * u < KH) {
                const float4 av = *reinterpret_cast<const float4 *>(ca + 4 * u);
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av.x, bq[u].x, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av.y, bq[u].y, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av.z, bq[u].z, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av.w, bq[u].w, acc, 0, 0, 0);
            }
        }
    }
    // four consecutive balls per group, read as one float4 each of norms,
    // their roots and radii
#pragma unroll
    for (int g4 = 0; g4 < 4; ++g4) {
        const int j = 8 * g4 + 4 * h;
        const float4 n4 = *reinterpret_cast<const float4 *>(tnrm + j);
        const float4 s4 = *reinterpret_cast<const float4 *>(tsq + j);
        const float4 r4 = *reinterpret_cast<const float4 *>(trad + j);
        out[4 * g4 + 0] = ball_lower_bound(pq, n4.x, s4.x, acc[4 * g4 + 0], r4.x, g15);
        out[4 * g4 + 1] = ball_lower_bound(pq, n4.y, s4.y, acc[4 * g4 + 1], r4.y, g15);
        out[4 * g4 + 2] = ball_lower_bound(pq, n4.z, s4.z, acc[4 * g4 + 2], r4.z, g15);
        out[4 * g4 + 3] = ball_lower_bound(pq, n4.w, s4.w, acc[4 * g4 + 3], r4.w, g15);
    }
}

// The ball-range split of the bounds grid: tiles per workgroup and groups.
__host__ __device__ inline int ball_ngrp(int nf, int tpw) { return ((nf + 31) / 32 + tpw - 1) / tpw; }
static int ball_tpw(int Q, int nf) {
    const long long qgs = (Q + 127) / 128, ntf = (nf + 31) / 32;
    return (int)std::max(1ll, std::min<long long>(kBallTilesPerWg, qgs * ntf / kBallMinWgs));
}

template <int KHM, int NCH>
__global__ __launch_bounds__(256, 4) void ball_lb_reg_kernel(const float *__restrict__ qpm, int pm_ld,
                                                             int Q, int dim,
                                                             const float *__restrict__ qnorm,
                                                             const float *__restrict__ fcent, int nf,
                                                             const float *__restrict__ fnorm,
                                                             const float *__restrict__ fsq,
                                                             const float *__restrict__ fradf,
                                                             const int *__restrict__ block_s,
                                                             const int *__restrict__ fb_start,
                                                             float *__restrict__ lb, int nb, int tpw,
                                                             const unsigned short *__restrict__ cbf,
                                                             float g15_bf) {
    __shared__ __align__(16) BallStage st;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 31, h = lane >> 5;
    const int ngrp = ball_ngrp(nf, tpw);
    const int qg = blockIdx.x / ngrp, grp = blockIdx.x - qg * ngrp;
    const int F0 = grp * tpw * 32;
    const int F1 = min(nf, F0 + tpw * 32);
    ball_stage(st, F0, F1, tpw, fnorm, fsq, fradf, block_s, fb_start);
    __syncthreads();
    const int q0 = qg * 128 + wave * 32;
    if (q0 >= Q) return;  // whole wave past the last query (no workgroup barrier below)
    const int qa = min(q0 + col, Q - 1);
    float4 bq[NCH ? 1 : KHM / 4];
    bf16x8_t bh[NCH ? NCH : 1], bl[NCH ? NCH : 1];
    ball_query_ops<KHM, NCH>(qpm, pm_ld, qa, dim, h, bq, bh, bl);
    const float g15 = NCH > 0 ? g15_bf : ball_error_coeff(dim);
    const float nq = qnorm[qa];
    const float pq = ball_norm_share(nq, nextafterf(sqrtf(nq), INFINITY), g15);
    const int q = q0 + col;
    // every lane runs the fold, so the open block (cur, its ball range) stays
    // wave-uniform in scalar registers and the row loop branches on scalars;
    // only the lanes of half 0 (one query each) store
    const bool folder = h == 0 && q < Q;
    int cur = -1, cur_s = 0, cur_e = 0;
    float m = INFINITY;
    auto flush = [&]() {
        if (cur < 0) return;
        const bool whole = cur_s >= F0 && cur_e <= F1;
        if (folder) {
            float *dst = lb + lb_at(cur, q, nb);
            if (whole)
                *dst = m;
            else
                atomicMin(reinterpret_cast<unsigned *>(dst), __float_as_uint(m));
        }
    };
    for (int tt = 0; tt < tpw; ++tt) {
        const int f0 = F0 + tt * 32;
        if (f0 >= F1) break;  // uniform over the wave
        float out[16];
        ball_tile_bounds<KHM, NCH>(st.nrm + (f0 - F0), st.sq + (f0 - F0), st.rad + (f0 - F0), f0, nf, pm_ld, dim,
                                   col, h, fcent, cbf, bq, bh, bl, pq, g15, out);
        // the partner half's rows (a lane swap across the halves), and
        // (lane r < 32) row r's block and that block's first ball and one
        // past its last
        float oth[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(out[i]), __float_as_uint(out[i]),
                                                             false, false);
            oth[i] = __uint_as_float(h ? sw[0] : sw[1]);
        }
        const int n = min(32, F1 - f0);
        const int rj = f0 - F0 + col;
        const int rb = st.blk[rj], rs = st.bs[rj], re = st.be[rj];
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            if (r >= n) continue;  // (not break: the loop must unroll)
            const int b = __builtin_amdgcn_readlane(rb, r);
            // on lane half 0, rows with (r / 4) even are its own
            const float v = ((r >> 2) & 1) ? oth[4 * (r >> 3) + (r & 3)] : out[4 * (r >> 3) + (r & 3)];
            if (b != cur) {
                flush();
                cur = b;
                cur_s = __builtin_amdgcn_readlane(rs, r);
                cur_e = __builtin_amdgcn_readlane(re, r);
                m = v;
            } else {
                m = fminf(m, v);
            }
        }
    }
    flush();
}

// Slot form of the same bounds.  Every block takes the smallest of 4, 8, 16
// or 32 consecutive slots that holds its balls (slots past its balls are
// empty balls: radius -1, bound +inf), blocks by descending slot count, so
// every block starts at a multiple of its size and never crosses a 32-slot
// tile.  In the accumulator of a tile, slots 8j + 4h .. 8j + 4h + 3 (group
// (j, h)) are the rows 4j .. 4j + 3 of lane half h, so one lane holds a whole
// 4-slot block: a block's bound is a min over one lane's four rows, over the
// two halves (8 slots: one lane swap), over two groups of both halves (16)
// or over the tile (32).  A tile's mode word holds 2 bits per j (0: the
// groups (j, 0) and (j, 1) are two 4-slot blocks, 1: one 8-slot block, 2: a
// 16-slot block over j and j ^ 1, 3: one 32-slot block) and blk[8t + 2j + h]
// the block that group (j, h) stores (-1: none, or not its first group).
// Each block is stored once per query, so the matrix needs no +inf fill
// and no atomics; the bounds equal the per-ball fold's bit for bit (the same
// products and the same per-ball formula, then an exact min).
constexpr int kSlotTilesPerWg = 16;  // slot tiles per workgroup for many queries
constexpr int kSlotWaves = 8;        // waves (of 32 queries) per workgroup
struct SlotStage {
    float pc[kSlotTilesPerWg * 32], rad[kSlotTilesPerWg * 32];  // (the slots' norm shares, radii)
    int blk[kSlotTilesPerWg * 8];
    int mode[kSlotTilesPerWg];
};

// Per-slot copies of the per-ball arrays: centres (rows of pm_ld floats),
// squared norms, their roots and the radii rounded up; an empty slot or
// empty ball has a zero centre and radius -inf, which the bound formula
// turns into +inf without a branch.
__global__ void ball_slot_gather_kernel(const int *__restrict__ sball, int nsl, const float *__restrict__ fcent,
                                        int pm_ld, const float *__restrict__ fnorm, const float *__restrict__ fsq,
                                        const float *__restrict__ fradf, float *__restrict__ scent,
                                        float *__restrict__ snrm, float *__restrict__ ssq, float *__restrict__ srad) {
    const long long n = (long long)nsl * pm_ld;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const int sl = (int)(e / pm_ld), d = (int)(e - (long long)sl * pm_ld);
        const int f = sball[sl];
        scent[e] = f >= 0 ? fcent[(long long)f * pm_ld + d] : 0.0f;
        if (d == 0) {
            snrm[sl] = f >= 0 ? fnorm[f] : 0.0f;
            ssq[sl] = f >= 0 ? fsq[f] : 0.0f;
            srad[sl] = f >= 0 && fradf[f] >= 0.0f ? fradf[f] : -INFINITY;
        }
    }
}

// Workgroup barrier for LDS hand-offs that waits for this wave's LDS
// operations only (global loads in flight keep running).
__device__ __forceinline__ void slot_lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// A workgroup: kSlotWaves waves of 32 queries x `tpw` consecutive slot
// tiles, as ball_lb_reg_kernel; the tiles' per-slot data and store map are staged
// in LDS.  Query columns past Q (up to Q rounded up to 64, the matrix's row
// groups) store +inf, so every entry of the matrix is written.
// NCH > 0: the workgroup also stages each tile's bfloat16 centres (32 slots
// x NCH x 64 bytes) in LDS, loaded once for all its waves: two buffers, one
// barrier per tile.  At the top of iteration tt the barrier makes tile tt
// (stored into buffer tt & 1 during iteration tt - 1, or before the loop)
// visible, and it follows every wave's reads of buffer (tt + 1) & 1 (made in
// iteration tt - 1), which iteration tt then overwrites with tile tt + 1.
// Every wave runs every iteration (waves past the last query compute the
// last query's bounds and store nothing), so every barrier is reached by all.
template <int KHM, int NCH>
__global__ __launch_bounds__(64 * kSlotWaves) void ball_lb_slot_kernel(const float *__restrict__ qpm, int pm_ld, int Q, int dim,
                                                           const float *__restrict__ qnorm,
                                                           const float *__restrict__ scent,
                                                           const unsigned short *__restrict__ sbf, int nsl,
                                                           const float *__restrict__ snrm,
                                                           const float *__restrict__ ssq,
                                                           const float *__restrict__ srad,
                                                           const int *__restrict__ sblk,
                                                           const int *__restrict__ smode, float *__restrict__ lb,
                                                           int nb, int tpw, float g15_bf) {
    constexpr int SLOT_B = NCH * 64 + 16;     // slot stride of the staged centres (bytes, padded)
    constexpr int NPIECE = NCH * 128;         // 16-byte pieces of one tile's centres
    constexpr int NT = 64 * kSlotWaves;       // threads
    constexpr int PPT = (NPIECE + NT - 1) / NT; // pieces per thread
    __shared__ __align__(16) SlotStage st;
    __shared__ __align__(16) unsigned char abuf[NCH ? 2 : 1][NCH ? 32 * SLOT_B : 16];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 31, h = lane >> 5;
    const int ntile = nsl / 32;
    const int ngrp = (ntile + tpw - 1) / tpw;
    const int qg = blockIdx.x / ngrp, grp = blockIdx.x - qg * ngrp;
    const int T0 = grp * tpw, nt = min(ntile, T0 + tpw) - T0;
    const float g15 = NCH > 0 ? g15_bf : ball_error_coeff(dim);
    for (int j = threadIdx.x; j < nt * 32; j += blockDim.x) {
        st.pc[j] = ball_norm_share(snrm[T0 * 32 + j], ssq[T0 * 32 + j], g15);
        st.rad[j] = srad[T0 * 32 + j];
    }
    for (int j = threadIdx.x; j < nt * 8; j += blockDim.x) st.blk[j] = sblk[T0 * 8 + j];
    if (threadIdx.x < nt) st.mode[threadIdx.x] = smode[T0 + threadIdx.x];
    const int q0 = qg * (32 * kSlotWaves) + wave * 32;
    const bool live = q0 < ((Q + 63) & ~63);  // wave-uniform: the wave's columns are rows of the matrix
    const int q = q0 + col;
    const int qa = min(q, Q - 1);
    float4 bq[NCH ? 1 : KHM / 4];
    bf16x8_t bh[NCH ? NCH : 1], bl[NCH ? NCH : 1];
    ball_query_ops<KHM, NCH>(qpm, pm_ld, qa, dim, h, bq, bh, bl);
    const float nq = qnorm[qa];
    const float pq = ball_norm_share(nq, nextafterf(sqrtf(nq), INFINITY), g15);
    const bool real = q < Q;
    constexpr int NA = NCH ? 2 * NCH : KHM / 4;
    const int KH = (dim + 1) / 2;
    u32x4_t av[NA];
    // NCH > 0: tile pieces in flight (registers) and their LDS stores
    u32x4_t pf[NCH ? PPT : 1];
    const unsigned char *sbytes = reinterpret_cast<const unsigned char *>(sbf);
    auto fetch = [&](int tile) {
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            const int pc = threadIdx.x + NT * i;
            if (pc < NPIECE)
                pf[i] = *reinterpret_cast<const u32x4_t *>(sbytes + ((long long)tile * NPIECE + pc) * 16);
        }
    };
    auto put = [&](int b) {
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            const int pc = threadIdx.x + NT * i;
            if (pc < NPIECE) {
                const int sl = pc / (NCH * 4), w = pc - sl * (NCH * 4);
                *reinterpret_cast<u32x4_t *>(&abuf[b][sl * SLOT_B + w * 16]) = pf[i];
            }
        }
    };
    // FP32 products: A operands from global memory, one tile ahead
    auto load_a = [&](int f0) {
        const float *ca = scent + (long long)(f0 + col) * pm_ld + h * KH;
#pragma unroll
        for (int u = 0; u < NA; ++u)
            av[u] = 4 * u < KH ? *reinterpret_cast<const u32x4_t *>(ca + 4 * u) : u32x4_t{0, 0, 0, 0};
    };
    if constexpr (NCH > 0) {
        if (nt > 0) {
            fetch(T0);
            put(0);
        }
    } else {
        if (nt > 0) load_a(T0 * 32);
    }
    __syncthreads();  // the staged slot data (and tile 0's centres)
    for (int tt = 0; tt < nt; ++tt) {
        f32x16_t acc = {};
        if constexpr (NCH > 0) {
            if (tt > 0) slot_lds_barrier();
            if (tt + 1 < nt) fetch(T0 + tt + 1);
            const unsigned char *ab = &abuf[tt & 1][col * SLOT_B + 16 * h];
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                av[2 * c] = *reinterpret_cast<const u32x4_t *>(ab + 64 * c);
                av[2 * c + 1] = *reinterpret_cast<const u32x4_t *>(ab + 64 * c + 32);
            }
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const bf16x8_t ah = __builtin_bit_cast(bf16x8_t, av[2 * c]);
                const bf16x8_t al = __builtin_bit_cast(bf16x8_t, av[2 * c + 1]);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh[c], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl[c], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh[c], acc, 0, 0, 0);
            }
            if (tt + 1 < nt) put((tt + 1) & 1);
        } else {
#pragma unroll
            for (int u = 0; u < NA; ++u) {
                if (4 * u < KH) {
                    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(av[u][0]), bq[u].x, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(av[u][1]), bq[u].y, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(av[u][2]), bq[u].z, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(av[u][3]), bq[u].w, acc, 0, 0, 0);
                }
            }
            if (tt + 1 < nt) load_a((T0 + tt + 1) * 32);
        }
        // the raw bounds of the tile's 32 slots (rows 8(i/4) + 4h + i%4), made
        // final after the fold; empty slots (radius -inf) come out +inf
        float out[16];
#pragma unroll
        for (int g4 = 0; g4 < 4; ++g4) {
            const int j = 32 * tt + 8 * g4 + 4 * h;
            const float4 c4 = *reinterpret_cast<const float4 *>(st.pc + j);
            const float4 r4 = *reinterpret_cast<const float4 *>(st.rad + j);
            out[4 * g4 + 0] = ball_bound_raw(pq, c4.x, acc[4 * g4 + 0], r4.x);
            out[4 * g4 + 1] = ball_bound_raw(pq, c4.y, acc[4 * g4 + 1], r4.y);
            out[4 * g4 + 2] = ball_bound_raw(pq, c4.z, acc[4 * g4 + 2], r4.z);
            out[4 * g4 + 3] = ball_bound_raw(pq, c4.w, acc[4 * g4 + 3], r4.w);
        }
        // group minima (one lane's four rows), the partner half's, and the
        // minima over 8, 16 and 32 slots
        float gm[4], m8[4];
#pragma unroll
        for (int g4 = 0; g4 < 4; ++g4) {
            gm[g4] = fminf(fminf(out[4 * g4], out[4 * g4 + 1]), fminf(out[4 * g4 + 2], out[4 * g4 + 3]));
            const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(gm[g4]), __float_as_uint(gm[g4]),
                                                             false, false);
            m8[g4] = fminf(gm[g4], __uint_as_float(h ? sw[0] : sw[1]));
        }
        const float m32 = fminf(fminf(m8[0], m8[1]), fminf(m8[2], m8[3]));
        const int mode = __builtin_amdgcn_readfirstlane(st.mode[tt]);
#pragma unroll
        for (int g4 = 0; g4 < 4; ++g4) {
            const int md = (mode >> (2 * g4)) & 3;
            const float v = md == 0 ? gm[g4] : md == 1 ? m8[g4] : md == 2 ? fminf(m8[g4], m8[g4 ^ 1]) : m32;
            const int b = st.blk[8 * tt + 2 * g4 + h];
            if (live && b >= 0) lb[lb_at(b, q, nb)] = real ? ball_bound_final(v) : INFINITY;
        }
    }
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
hipError_t idx_gather_pack(const float *pm_src, int pm_ld, int dim, int oct, const int *perm,
                           int n_pos, int n_pad, float *packed, float *pm_dst, float *norms,
                           hipStream_t s) {
    hipLaunchKernelGGL(gather_pack_kernel, dim3((n_pad + 255) / 256), dim3(256), 0, s, pm_src,
                       pm_ld, dim, oct, perm, n_pos, n_pad, packed, pm_dst, norms);
    return hipGetLastError();
}
hipError_t idx_pick_blocks(const float *score, int Q, int nb, int P, const float *lb,
                           const int *counts, int *chosen, float *lb_rest, hipStream_t s) {
    if (Q <= kPickSingleMaxQ) {
        const size_t lds = (size_t)nb * sizeof(float);
        if (lds > 65536) {
            hipError_t e = hipFuncSetAttribute((const void *)pick_blocks_single_kernel,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(pick_blocks_single_kernel, dim3(Q), dim3(256), lds, s, score, Q, nb, P, lb,
                           counts, chosen, lb_rest);
    } else if (P <= 8 && score)
        hipLaunchKernelGGL((pick_blocks_list_kernel<8, true>), dim3((Q + 63) / 64), dim3(64 * kPickWaves), 0, s,
                           score, Q, nb, P, lb, counts, chosen, lb_rest);
    else if (P <= 8 && nb <= kPickKeyB)  // (the no-score keys hold block numbers below 2^30)
        hipLaunchKernelGGL((pick_blocks_list_kernel<8, false>), dim3((Q + 63) / 64), dim3(64 * kPickWaves), 0, s,
                           score, Q, nb, P, lb, counts, chosen, lb_rest);
    else if (P <= 16 && score)
        hipLaunchKernelGGL((pick_blocks_list_kernel<16, true>), dim3((Q + 63) / 64), dim3(64 * kPickWaves), 0,
                           s, score, Q, nb, P, lb, counts, chosen, lb_rest);
    else if (P <= 16 && nb <= kPickKeyB)
        hipLaunchKernelGGL((pick_blocks_list_kernel<16, false>), dim3((Q + 63) / 64), dim3(64 * kPickWaves), 0,
                           s, score, Q, nb, P, lb, counts, chosen, lb_rest);
    else
        hipLaunchKernelGGL(pick_blocks_kernel, dim3((Q + 63) / 64), dim3(64 * kPickWaves), 0, s,
                           score, Q, nb, P, lb, counts, chosen, lb_rest);
    return hipGetLastError();
}
hipError_t idx_router_first(const float *score, int Q, int nb, const int *counts, float *lb,
                            hipStream_t s) {
    hipLaunchKernelGGL(router_first_kernel, dim3((Q + 63) / 64), dim3(64 * kPickWaves), 0, s, score, Q,
                       nb, counts, lb);
    return hipGetLastError();
}
hipError_t idx_pair_count(const int *chosen, int Q, int P, int p0, int p1, int nu, int *counts,
                          hipStream_t s) {
    hipLaunchKernelGGL(pair_count_kernel, dim3(grid_for((long long)Q * (p1 - p0))), dim3(256), 0, s,
                       chosen, Q, P, p0, p1, nu, counts);
    return hipGetLastError();
}
hipError_t idx_pair_scatter(const int *chosen, int Q, int P, int p0, int p1, int nu, int *cursor,
                            int *qlist, int *rows, int row_base, hipStream_t s) {
    hipLaunchKernelGGL(pair_scatter_kernel, dim3(grid_for((long long)Q * (p1 - p0))), dim3(256), 0,
                       s, chosen, Q, P, p0, p1, nu, cursor, qlist, rows, row_base);
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
hipError_t idx_unit_tasks(const int *counts, int nu, int qpt, const int *unit_start, const int *unit_count,
                          int *qstart, int *cursor, ScanTask *tasks, int max_tasks, int *clear, hipStream_t s) {
    if (qpt < 1 || qpt > kQueriesPerWave) return hipErrorInvalidValue;
    hipLaunchKernelGGL(unit_tasks_kernel, dim3(kUnitTaskWgs), dim3(kUnitTaskThreads), 0, s, counts, nu, qpt,
                       unit_start, unit_count, qstart, cursor, tasks, max_tasks, clear);
    return hipGetLastError();
}
hipError_t idx_ref_block(const int *block_start, const int *block_count, int nb, const int *perm,
                         int *out, hipStream_t s) {
    hipLaunchKernelGGL(ref_block_kernel, dim3(nb), dim3(256), 0, s, block_start, block_count, perm,
                       out);
    return hipGetLastError();
}

hipError_t idx_row_norms(const float *pm, int ld, int n, int dim, float *out, hipStream_t s) {
    hipLaunchKernelGGL(row_norms_kernel, dim3(grid_for(n)), dim3(256), 0, s, pm, ld, n, dim, out);
    return hipGetLastError();
}

hipError_t idx_kpp_seed(const float *pm, int ld, int dim, long long n, int m, int nb,
                        unsigned long long seed, float *mind, int *cidx, double *wsum,
                        unsigned *ticket, float *cent, hipStream_t s) {
    const int nwg = std::max(1, std::min(kKppGroups, (m + 255) / 256));
    hipLaunchKernelGGL(kpp_init_kernel, dim3(grid_for(m)), dim3(256), 0, s, mind, m, cidx, ticket,
                       seed);
    for (int t = 1; t < nb; ++t)
        hipLaunchKernelGGL(kpp_step_kernel, dim3(nwg), dim3(256), (size_t)ld * 4, s, pm, ld, dim, n,
                           m, mind, cidx, t, seed, wsum, ticket);
    hipLaunchKernelGGL(kpp_gather_kernel, dim3(grid_for((long long)nb * ld)), dim3(256), 0, s, pm,
                       ld, n, m, cidx, nb, cent);
    return hipGetLastError();
}
hipError_t idx_group_fine(const int *cand, int kc, int nf, const int *fcount, int nb, int cap,
                          int *fblock, hipStream_t s) {
    const size_t lds = (size_t)nb * 4;
    if (lds > 65536) {
        const hipError_t e = hipFuncSetAttribute((const void *)group_fine_kernel,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(group_fine_kernel, dim3(1), dim3(64), lds, s, cand, kc, nf, fcount, nb, cap,
                       fblock);
    return hipGetLastError();
}
hipError_t idx_block_count(const int *fblock, const int *fcount, int nf, int *bcount, int *bfine,
                           hipStream_t s) {
    hipLaunchKernelGGL(block_count_kernel, dim3(grid_for(nf)), dim3(256), 0, s, fblock, fcount, nf,
                       bcount, bfine);
    return hipGetLastError();
}
hipError_t idx_fine_start(const int *fblock, const int *fcount, int nf, int *pos_cursor,
                          int *fstart, int *slot_cursor, int *fine_sorted, hipStream_t s) {
    hipLaunchKernelGGL(fine_start_kernel, dim3(grid_for(nf)), dim3(256), 0, s, fblock, fcount, nf,
                       pos_cursor, fstart, slot_cursor, fine_sorted);
    return hipGetLastError();
}
hipError_t idx_finalize_fine(const float *pm, int ld, int dim, const float *fcent,
                             const int *fine_sorted, const int *fstart, const int *fcount,
                             const int *fblock, int nf, float *cent_s, double *rad_s, int *block_s,
                             hipStream_t s) {
    hipLaunchKernelGGL(finalize_fine_kernel, dim3(nf), dim3(256), 0, s, pm, ld, dim, fcent,
                       fine_sorted, fstart, fcount, fblock, cent_s, rad_s, block_s);
    return hipGetLastError();
}
hipError_t idx_ball_refit(const float *pm, int ld, int dim, const int *block_start, const int *block_count,
                          const int *fb_start, int nb, const float *fcent, int nf,
                          unsigned long long *d2bits, int *member, double *rad, hipStream_t s) {
    hipError_t e = hipMemsetAsync(d2bits, 0, (size_t)nf * sizeof(unsigned long long), s);
    if (e == hipSuccess) e = hipMemsetAsync(member, 0, (size_t)nf * sizeof(int), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(ball_refit_kernel, dim3(nb), dim3(256), 0, s, pm, ld, dim, block_start, block_count,
                       fb_start, fcent, d2bits, member);
    hipLaunchKernelGGL(ball_refit_finish_kernel, dim3(grid_for(nf)), dim3(256), 0, s, d2bits, member, nf,
                       rad);
    return hipGetLastError();
}
hipError_t idx_ball_first(const float *qpm, int pm_ld, int dim, int Q, const int *chosen, int P,
                          const float *fcent, const double *frad, const int *fb_start, int *out, int Pb,
                          hipStream_t s) {
    hipLaunchKernelGGL(ball_first_kernel, dim3((Q + kBallWaves - 1) / kBallWaves), dim3(64 * kBallWaves), 0, s,
                       qpm, pm_ld, dim, Q, chosen, P, fcent, frad, fb_start, out, Pb);
    return hipGetLastError();
}
int idx_ball_cand_max() { return kBallCand; }
hipError_t idx_ball_select(const float *qpm, int pm_ld, int dim, int Q, int kp, const int *chosen, int P,
                           const float *fcent, const double *frad, const int *fcnt, const int *fb_start, int *out,
                           int Pb, int ncand, float *lb_rest, hipStream_t s, const float *qnorm,
                           const unsigned *rmax, double eq, double er, unsigned *theta) {
    const int ncap = (ncand + 7) / 8 * 8;
    if (ncap < 8 || ncap > kBallCand) return hipErrorInvalidValue;
    if (ncand <= kSelLaneCand * kWave && P <= kWave) {
        hipLaunchKernelGGL(ball_select_lanes_kernel, dim3((Q + kBallWaves - 1) / kBallWaves), dim3(64 * kBallWaves),
                           0, s, qpm, pm_ld, dim, Q, kp, chosen, P, fcent, frad, fcnt, fb_start, out, Pb, lb_rest,
                           qnorm, rmax, eq, er, theta);
        return hipGetLastError();
    }
    if (theta) {  // (this form seeds nothing: every threshold starts at +inf)
        const hipError_t e = hipMemsetD32Async((hipDeviceptr_t)theta, 0xFF800000u, Q, s);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(ball_select_kernel, dim3((Q + kBallWaves - 1) / kBallWaves), dim3(64 * kBallWaves),
                       (size_t)kBallWaves * ncap * kSelBytes, s, qpm, pm_ld, dim, Q, kp, chosen, P, fcent, frad, fcnt,
                       fb_start, out, Pb, lb_rest, ncap);
    return hipGetLastError();
}
hipError_t idx_ball_rest(const float *qpm, int pm_ld, int dim, int Q, const float *qnorm, const unsigned *theta,
                         const unsigned *rmax, double eq, double er, int res, const int *chosen, int P,
                         const float *fcent, const double *frad, const int *fb_start, int *out, int Pb, int cap,
                         float *lb_rest, hipStream_t s) {
    if (cap < 0 || cap > Pb - 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(ball_rest_kernel, dim3((Q + kBallWaves - 1) / kBallWaves), dim3(64 * kBallWaves), 0, s,
                       qpm, pm_ld, dim, Q, qnorm, theta, rmax, eq, er, res, chosen, P, fcent, frad, fb_start, out,
                       Pb, cap, lb_rest);
    return hipGetLastError();
}
hipError_t idx_ball_range_radius(const float *pm, int ld, int dim, const float *fcent, const int *ball_pos,
                                 const int *ball_cnt, int nf, double *rad, hipStream_t s) {
    hipLaunchKernelGGL(ball_range_radius_kernel, dim3(nf), dim3(256), 0, s, pm, ld, dim, fcent, ball_pos, ball_cnt,
                       rad);
    return hipGetLastError();
}
hipError_t idx_ball_aux(const float *fnorm, const double *frad, int nf, float *fsq, float *fradf,
                        hipStream_t s) {
    hipLaunchKernelGGL(ball_aux_kernel, dim3(grid_for(nf)), dim3(256), 0, s, fnorm, frad, nf, fsq, fradf);
    return hipGetLastError();
}
hipError_t idx_pack_ball_bf3(const float *fcent, int pm_ld, int nf, int dim, int nch, unsigned short *out,
                             hipStream_t s) {
    hipLaunchKernelGGL(ball_bf3_pack_kernel, dim3(grid_for((long long)nf * nch * 16)), dim3(256), 0, s,
                       fcent, pm_ld, nf, dim, nch, out);
    return hipGetLastError();
}

// The tiled bound matrix (lb_at) as nb rows of Q: out[b * Q + q].
__global__ void lb_untile_kernel(const float *__restrict__ tiled, int Q, int nb, float *__restrict__ out) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < (long long)Q * nb;
         e += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(e / Q), q = (int)(e - (long long)b * Q);
        out[e] = tiled[lb_at(b, q, nb)];
    }
}
hipError_t idx_lb_untile(const float *tiled, int Q, int nb, float *out, hipStream_t s) {
    hipLaunchKernelGGL(lb_untile_kernel, dim3(grid_for((long long)Q * nb)), dim3(256), 0, s, tiled, Q, nb, out);
    return hipGetLastError();
}
hipError_t idx_ball_lb(const float *qpm, int pm_ld, int Q, int dim, const float *qnorm,
                       const float *fcent, int nf, const float *fnorm, const float *fsq,
                       const float *fradf, const int *block_s, const int *fb_start, int nb, float *lb,
                       hipStream_t s, const unsigned short *cbf, int nch_bf, float g15_bf) {
    hipError_t e = hipMemsetD32Async((hipDeviceptr_t)lb, 0x7F800000, (size_t)lb_elems(Q, nb), s);
    if (e != hipSuccess) return e;
    const int KH = (dim + 1) / 2;
    const long long qgs = (Q + 127) / 128;
    const int tpw = ball_tpw(Q, nf);
    const long long wgs = qgs * ball_ngrp(nf, tpw);
    if (cbf && nch_bf >= 1 && nch_bf <= 8 && 16 * nch_bf >= dim) {
#define BF(N)                                                                                          \
    if (nch_bf == N) {                                                                                 \
        hipLaunchKernelGGL((ball_lb_reg_kernel<4, N>), dim3((unsigned)wgs), dim3(256), 0, s, qpm, pm_ld, Q, \
                           dim, qnorm, fcent, nf, fnorm, fsq, fradf, block_s, fb_start, lb, nb, tpw, cbf, \
                           g15_bf);                                                                    \
        return hipGetLastError();                                                                      \
    }
        BF(1) BF(2) BF(3) BF(4) BF(5) BF(6) BF(7) BF(8)
#undef BF
    }
    if (KH % 4 == 0 && pm_ld % 4 == 0 && KH <= 64) {
        if (KH <= 32)
            hipLaunchKernelGGL((ball_lb_reg_kernel<32, 0>), dim3((unsigned)wgs), dim3(256), 0, s, qpm, pm_ld,
                               Q, dim, qnorm, fcent, nf, fnorm, fsq, fradf, block_s, fb_start, lb, nb, tpw,
                               nullptr, 0.0f);
        else
            hipLaunchKernelGGL((ball_lb_reg_kernel<64, 0>), dim3((unsigned)wgs), dim3(256), 0, s, qpm, pm_ld,
                               Q, dim, qnorm, fcent, nf, fnorm, fsq, fradf, block_s, fb_start, lb, nb, tpw,
                               nullptr, 0.0f);
        return hipGetLastError();
    }
    const long long tiles = (long long)((Q + 31) / 32) * ((nf + 31) / 32);
    hipLaunchKernelGGL(ball_lb_kernel, dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, s, qpm,
                       pm_ld, Q, dim, qnorm, fcent, nf, fnorm, fsq, fradf, block_s, fb_start, lb, nb);
    return hipGetLastError();
}
hipError_t idx_ball_slot_gather(const int *sball, int nsl, const float *fcent, int pm_ld, const float *fnorm,
                                const float *fsq, const float *fradf, float *scent, float *snrm, float *ssq,
                                float *srad, hipStream_t s) {
    hipLaunchKernelGGL(ball_slot_gather_kernel, dim3(grid_for((long long)nsl * pm_ld)), dim3(256), 0, s, sball, nsl,
                       fcent, pm_ld, fnorm, fsq, fradf, scent, snrm, ssq, srad);
    return hipGetLastError();
}
bool idx_ball_lb_slots_fit(int dim, int pm_ld, int nch_bf) {
    if (nch_bf) return nch_bf >= 1 && nch_bf <= 8 && 16 * nch_bf >= dim;
    const int KH = (dim + 1) / 2;
    return KH % 4 == 0 && pm_ld % 4 == 0 && KH <= 64;
}
hipError_t idx_ball_lb_slots(const float *qpm, int pm_ld, int Q, int dim, const float *qnorm, const float *scent,
                             const unsigned short *sbf, int nch_bf, float g15_bf, int nsl, const float *snrm,
                             const float *ssq, const float *srad, const int *sblk, const int *smode, int nb,
                             float *lb, hipStream_t s) {
    if (nsl < 32 || nsl % 32 != 0 || !idx_ball_lb_slots_fit(dim, pm_ld, nch_bf)) return hipErrorInvalidValue;
    const long long qgs = (Q + 32 * kSlotWaves - 1) / (32 * kSlotWaves), ntile = nsl / 32;
    const int tpw = (int)std::max(1ll, std::min<long long>(kSlotTilesPerWg, qgs * ntile / kBallMinWgs));
    const long long wgs = qgs * ((ntile + tpw - 1) / tpw);
#define SL(KHM, N)                                                                                               \
    hipLaunchKernelGGL((ball_lb_slot_kernel<KHM, N>), dim3((unsigned)wgs), dim3(64 * kSlotWaves), 0, s, qpm, pm_ld, Q, dim,   \
                       qnorm, scent, sbf, nsl, snrm, ssq, srad, sblk, smode, lb, nb, tpw, g15_bf);               \
    return hipGetLastError();
    if (nch_bf) {
        switch (nch_bf) {
        case 1: SL(4, 1)
        case 2: SL(4, 2)
        case 3: SL(4, 3)
        case 4: SL(4, 4)
        case 5: SL(4, 5)
        case 6: SL(4, 6)
        case 7: SL(4, 7)
        default: SL(4, 8)
        }
    }
    if ((dim + 1) / 2 <= 32) {
        SL(32, 0)
    }
    SL(64, 0)
#undef SL
}

// Batched outputs: element e = j * Qtot + q of the class's k x Qtot results
// goes to row j of the problem whose queries hold column q (outs by
// ascending qstart), at its own column q - qstart.
__global__ void batch_outputs_kernel(const float *__restrict__ dist, const int *__restrict__ idx, int k,
                                     int Qtot, const BatchOut *__restrict__ outs, int np) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < (long long)k * Qtot;
         e += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(e / Qtot), q = (int)(e - (long long)j * Qtot);
        int lo = 0, hi = np - 1;  // the last problem with qstart <= q
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (outs[mid].qstart <= q)
                lo = mid;
            else
                hi = mid - 1;
        }
        const BatchOut o = outs[lo];
        const long long d = (long long)j * o.qn + (q - o.qstart);
        o.dist[d] = dist[e];
        o.idx[d] = idx[e];
    }
}
hipError_t idx_batch_outputs(const float *dist, const int *idx, int k, int Qtot, const BatchOut *outs,
                             int np, hipStream_t s) {
    hipLaunchKernelGGL(batch_outputs_kernel, dim3(grid_for((long long)k * Qtot)), dim3(256), 0, s, dist, idx, k,
                       Qtot, outs, np);
    return hipGetLastError();
}
hipError_t idx_batch_perm(const int *starts, const int *counts, int nsets, int n_pos, int *perm,
                          hipStream_t s) {
    hipLaunchKernelGGL(batch_perm_kernel, dim3(grid_for(n_pos)), dim3(256), 0, s, starts, counts,
                       nsets, n_pos, perm);
    return hipGetLastError();
}

}  // namespace knnhip
