// Stages around the scan: squared norms, candidate selection with exact FP32
// re-rank and certificate, and the exact per-query full scan used when a
// certificate fails.
//
// Exactness argument (the result equals the reference ground truth
// code/test.cpp:129-175 bit for bit):
//   * the scan leaves, per query, a candidate set C and a threshold tau such
//     that every reference outside C has approximate key >= tau;
//   * the approximate key a(r) = ||r||^2 - 2 r.q (FP32 MFMA) differs from the
//     true ||r-q||^2 - ||q||^2 by at most g*(||q||^2 + 3.1*max||r||^2), with
//     g = 2*(dim+4)*2^-24 (twice the standard gamma bound of a dim+2 term
//     FP32 dot product and of the FP32 norm);
//   * the ground-truth FP32 sum differs from ||r-q||^2 by at most the
//     relative factor g;
//   * so every reference outside C has ground-truth sum >= L, computed in
//     double below.  If L > (k-th key * (1 + 2^-22))^2, every outside
//     reference has sqrtf key strictly above the k-th candidate key, and the
//     k best candidates under the reference's own (sqrtf key, index) order
//     are the reference's answer.  Otherwise the query is answered by the
//     exact full scan.
#include <limits.h>

#include "knn_common.h"

namespace knnhip {

// ---------------------------------------------------------------------------
// Select the kp smallest approximate keys of the union of the slice heaps,
// re-rank them with the ground-truth FP32 distance, write the k best and
// check the certificate.  One wave per query.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
// The same sum for lane values in [0, 2^B): one ballot per bit and scalar
// bit counts, no cross-lane shuffles (a bisection step's latency).
template <int B>
__device__ __forceinline__ int wave_sum_bits(int v) {
    int s = 0;
#pragma unroll
    for (int b = 0; b < B; ++b) s += __popcll(__ballot((v >> b) & 1)) << b;
    return s;
}

template <int NPL>
__global__ __launch_bounds__(256) void select_rerank_kernel(FinalArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int waves = blockDim.x >> 6;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int q = blockIdx.x * waves + wave;
    if (q >= a.Q) return;
    if (a.nonfinite && *a.nonfinite) {
        // non-finite input: no answer, and no fallback either
        for (int j = lane; j < a.k; j += kWave) {
            a.out_dist[(long long)j * a.Q + q] = __uint_as_float(0x7FC00000u);
            a.out_idx[(long long)j * a.Q + q] = -1;
        }
        return;
    }

    const int maxc = a.maxc;
    int *cidx = reinterpret_cast<int *>(smem) + (size_t)wave * maxc;      // reference position
    int *corg = reinterpret_cast<int *>(smem) + (size_t)(waves + wave) * maxc;  // caller's index
    float *ckey = reinterpret_cast<float *>(smem + (size_t)waves * maxc * 2 * sizeof(int)) +
                  (size_t)wave * maxc;
    float *kth_slot = reinterpret_cast<float *>(smem + (size_t)waves * maxc * 3 * sizeof(int)) + wave;

    // 1. Gather the union of the query's heaps (key order-mapped; placeholders
    //    of unfilled heaps map to the largest code and are never counted).
    const int kp = a.kp;
    const int U = a.nrows * kp;
    unsigned code[NPL];
    int ids[NPL];
    int nreal = 0;
    unsigned cmin = 0xFFFFFFFFu, cmax = 0u;  // the lane's smallest and largest real codes
#pragma unroll
    for (int j = 0; j < NPL; ++j) {
        const int e = j * kWave + lane;
        code[j] = 0xFFFFFFFFu;
        ids[j] = -1;
        if (e < U) {
            const int s = e / kp, t = e - s * kp;
            const long long row = a.rows ? (long long)a.rows[(long long)q * a.nrows + s]
                                         : s * a.row_stride + q;
            const int id = row >= 0 ? a.part_idx[row * kp + t] : -1;
            const long long off = row * kp + t;
            if (id >= 0) {
                ids[j] = id;
                code[j] = float_to_ordered(a.part_val[off]);
                cmin = min(cmin, code[j]);
                cmax = max(cmax, code[j]);
                ++nreal;
            }
        }
    }
    const int N = wave_sum(nreal);

    // 2. cut = kp-th smallest key (bisection on the order-mapped code); with
    //    fewer than kp real entries every real entry is a candidate.
    unsigned cut = 0xFFFFFFFEu;
    const bool all_in = N < kp;
    if (!all_in) {
        // the cut lies between the smallest and the largest real code (at
        // least kp real codes lie at or below the largest), so the bisection
        // starts from that range: about 20 steps instead of 32 for keys that
        // share an exponent
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            cmin = min(cmin, (unsigned)__shfl_xor((int)cmin, off));
            cmax = max(cmax, (unsigned)__shfl_xor((int)cmax, off));
        }
        unsigned lo = __builtin_amdgcn_readfirstlane(cmin), hi = __builtin_amdgcn_readfirstlane(cmax);
        while (lo < hi) {
            const unsigned mid = lo + ((hi - lo) >> 1);
            int c = 0;
#pragma unroll
            for (int j = 0; j < NPL; ++j) c += code[j] <= mid;
            // (a lane's count is at most NPL <= 32: six bits)
            if (wave_sum_bits<NPL < 32 ? 5 : 6>(c) >= kp)
                hi = mid;
            else
                lo = mid + 1;
        }
        cut = lo;
    }

    // 3. Compact the candidates (code <= cut) in a fixed order.
    int nc = 0;
#pragma unroll
    for (int j = 0; j < NPL; ++j) {
        const bool pred = code[j] <= cut;
        const unsigned long long m = __ballot(pred);
        const int pos = nc + __popcll(m & ((1ull << lane) - 1ull));
        if (pred && pos < maxc) cidx[pos] = ids[j];
        nc += __popcll(m);
    }
    const bool overflow = nc > maxc;
    wave_sync();

    // 4. Ground-truth FP32 distance of every candidate, then rank by the
    //    reference order (sqrtf key, index) and write the k best.
    if (!overflow && nc <= kWave) {
        // one candidate a lane: its distance in registers, its rank by reading
        // the others' (key, index) across the wave
        float kc = INFINITY;
        int ic = INT_MAX;
        if (lane < nc) {
            const int r = cidx[lane];
            kc = sqrt_rn(canonical_sq_rows(a.ref_pm + (long long)r * a.pm_ld, a.query_pm + (long long)q * a.pm_ld,
                                           a.dim));
            ic = a.perm ? a.perm[r] : r;
        }
        int rank = 0;
        for (int j = 0; j < nc; ++j) {
            const float y = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(kc), j));
            const int yi = __builtin_amdgcn_readlane(ic, j);
            rank += lex_less(y, yi, kc, ic);
        }
        if (lane < nc) {
            if (rank < a.k) {
                a.out_dist[(long long)rank * a.Q + q] = kc;
                a.out_idx[(long long)rank * a.Q + q] = ic;
            }
            if (rank == a.k - 1) *kth_slot = kc;
        }
        wave_sync();
    } else if (!overflow) {
        for (int c = lane; c < nc; c += kWave) {
            const int r = cidx[c];
            ckey[c] = sqrt_rn(canonical_sq_rows(a.ref_pm + (long long)r * a.pm_ld,
                                                a.query_pm + (long long)q * a.pm_ld, a.dim));
            corg[c] = a.perm ? a.perm[r] : r;
        }
        wave_sync();
        for (int c = lane; c < nc; c += kWave) {
            const float kc = ckey[c];
            const int ic = corg[c];
            int rank = 0;
            for (int j = 0; j < nc; ++j) rank += lex_less(ckey[j], corg[j], kc, ic);
            if (rank < a.k) {
                a.out_dist[(long long)rank * a.Q + q] = kc;
                a.out_idx[(long long)rank * a.Q + q] = ic;
            }
            if (rank == a.k - 1) *kth_slot = kc;
        }
        wave_sync();
    }

    // 5. Certificate.
    if (lane == 0) {
        bool ok = !overflow && !a.force_fallback && nc >= a.k;
        // Every reference outside the candidates has key >= tau: entries of
        // the union past the cut, and references the scan rejected against
        // a threshold no smaller than the final shared one.
        float tau_f = ordered_to_float(a.theta[q]);
        if (!all_in) tau_f = fminf(tau_f, ordered_to_float(cut));
        if (ok && tau_f < INFINITY) {
            const double tau = (double)tau_f;
            const double g = 2.0 * (double)(a.dim + 4) * 5.9604644775390625e-8;
            double lower;
            if (a.res_row) {
                // residual keys: a rejected reference r of the block of row j
                // has approximate squared distance >= tau, so
                // ||q - r|| >= max(L_j, sqrt(tau - eta_j) - eps_j); the
                // smallest such bound over the query's rows bounds every
                // rejected reference, and its ground-truth FP32 sum is within
                // the relative factor g of ||q - r||^2
                double bmin = INFINITY;
                for (int j = 0; j < a.nrows; ++j) {
                    const long long row = a.rows ? (long long)a.rows[(long long)q * a.nrows + j]
                                                 : j * a.row_stride + q;
                    if (row < 0) continue;
                    const float4 e = *reinterpret_cast<const float4 *>(a.res_row + 4 * row);
                    const double t2 = tau - (double)e.z;
                    const double dl = (t2 > 0.0 ? sqrt(t2) * (1.0 - 1e-15) : 0.0) - (double)e.y;
                    bmin = fmin(bmin, fmax((double)e.x, dl));
                }
                lower = bmin > 0.0 ? bmin * bmin * (1.0 - g) : 0.0;
            } else {
                const double nq = (double)a.qnorm[q];
                const double nr = (double)__uint_as_float(*a.rmax);
                const double eq = a.err_q > 0.0 ? a.err_q : g, er = a.err_r > 0.0 ? a.err_r : 3.1 * g;
                lower = (tau + nq * (1.0 - g) - (eq * nq + er * nr) - (double)a.dim * 1e-36) * (1.0 - g);
            }
            const double kth = (double)(*kth_slot) * (1.0 + 2.384185791015625e-7);
            ok = lower > kth * kth;
        }
        if (ok && a.lb_rest) {
            // References of skipped blocks are at least lb_rest away (triangle
            // inequality); their ground-truth sum is within the relative factor
            // g of the true squared distance.
            const double lb = (double)a.lb_rest[q];
            const double g = 2.0 * (double)(a.dim + 4) * 5.9604644775390625e-8;
            const double kth = (double)(*kth_slot) * (1.0 + 2.384185791015625e-7);
            ok = lb * lb * (1.0 - g) > kth * kth;
        }
        if (!ok) {
            const int slot = atomicAdd(a.fb_count, 1);
            a.fb_list[slot] = q;
        }
    }
}

// ---------------------------------------------------------------------------
// Exact per-query scan: the reference ground truth computed on the device.
// One wave per query; lane l takes references l, l+64, ... in increasing
// order and keeps its own top-k with the reference's insertion semantics
// (code/test.cpp:85-114); the 64 sorted lists are then merged in
// (key, index) order.  Used for queries whose certificate failed, and for
// every query when k is too large for the scan's heaps.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void exact_scan_kernel(ExactArgs a) {
    const int lane = threadIdx.x;
    const int n = a.count ? *a.count : a.Q;
    const int k = a.k;
    float *mykey = a.scratch_key + ((long long)blockIdx.x * kWave + lane) * k;
    int *myidx = a.scratch_idx + ((long long)blockIdx.x * kWave + lane) * k;
    const bool bad = a.nonfinite && *a.nonfinite;
    for (int t = blockIdx.x; t < n; t += gridDim.x) {
        const int q = a.list ? a.list[t] : t;
        if (bad) {
            // non-finite input (see FinalArgs::nonfinite)
            for (int j = lane; j < k; j += kWave) {
                a.out_dist[(long long)j * a.Q + q] = __uint_as_float(0x7FC00000u);
                a.out_idx[(long long)j * a.Q + q] = -1;
            }
            continue;
        }
        int filled = 0;
        for (int r = lane; r < a.R; r += kWave) {
            const float key = sqrt_rn(canonical_sq(a.ref, a.ldr, r, a.query, a.ldq, q, a.dim));
            if (filled == k && !(key < mykey[k - 1])) continue;
            int pos = filled < k ? filled : k - 1;
            while (pos > 0 && mykey[pos - 1] > key) {
                mykey[pos] = mykey[pos - 1];
                myidx[pos] = myidx[pos - 1];
                --pos;
            }
            mykey[pos] = key;
            myidx[pos] = r;
            if (filled < k) ++filled;
        }
        int head = 0;
        for (int j = 0; j < k; ++j) {
            const float hk = head < filled ? mykey[head] : INFINITY;
            const int hid = head < filled ? myidx[head] : INT_MAX;
            float bk = hk;
            int bi = hid;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                const float ok = __shfl_xor(bk, off);
                const int oi = __shfl_xor(bi, off);
                if (lex_less(ok, oi, bk, bi)) {
                    bk = ok;
                    bi = oi;
                }
            }
            if (head < filled && hid == bi) ++head;
            if (lane == 0) {
                a.out_dist[(long long)j * a.Q + q] = bk;
                a.out_idx[(long long)j * a.Q + q] = bi;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
size_t select_lds_bytes(int maxc, int waves) {
    return (size_t)waves * (maxc * 3 * sizeof(int) + sizeof(float)) + 16;
}

template <int NPL>
static hipError_t launch_select_t(const FinalArgs &a, hipStream_t s) {
    const int waves = 4;
    const size_t lds = select_lds_bytes(a.maxc, waves);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)select_rerank_kernel<NPL>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(select_rerank_kernel<NPL>, dim3((a.Q + waves - 1) / waves),
                       dim3(waves * kWave), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_select(const FinalArgs &a, hipStream_t s) {
    const int U = a.nrows * a.kp;
    if (U <= 4 * kWave) return launch_select_t<4>(a, s);
    if (U <= 8 * kWave) return launch_select_t<8>(a, s);
    if (U <= 16 * kWave) return launch_select_t<16>(a, s);
    if (U <= 32 * kWave) return launch_select_t<32>(a, s);
    return hipErrorInvalidValue;
}

hipError_t launch_exact(const ExactArgs &a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(exact_scan_kernel, dim3(grid), dim3(kWave), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Seeded generators (same hash as the oracle's CPU generator).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t rng_u64(uint64_t seed, uint64_t stream, uint64_t counter) {
    return mix64(seed * 0x9E3779B97F4A7C15ull + stream * 0xD1B54A32D192ED03ull +
                 counter * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull);
}
__device__ __forceinline__ float unit24(uint64_t u) {
    return (float)(u >> 40) * (1.0f / 16777216.0f);
}

__global__ void gen_uniform_kernel(float *out, int n, int dim, uint64_t seed, float scale) {
#pragma clang fp contract(off)
    const long long total = (long long)n * dim;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long d = e / n, i = e - d * n;
        out[e] = scale * unit24(rng_u64(seed, 0, (uint64_t)i * dim + d));
    }
}

__global__ void gen_mixture_kernel(float *out, int n, int dim, int ncomp, float sigma,
                                   uint64_t centre_seed, uint64_t point_seed) {
#pragma clang fp contract(off)
    const long long total = (long long)n * dim;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long d = e / n, i = e - d * n;
        const uint64_t c = ((rng_u64(point_seed, 2, (uint64_t)i) >> 32) * (uint64_t)ncomp) >> 32;
        const float centre =
            2.0f * unit24(rng_u64(centre_seed, 1, c * (uint64_t)dim + d)) - 1.0f;
        const uint64_t cnt = (uint64_t)i * dim + d;
        const double u1 = ((double)(rng_u64(point_seed, 3, cnt) >> 11) + 0.5) *
                          (1.0 / 9007199254740992.0);
        const double u2 =
            (double)(rng_u64(point_seed, 4, cnt) >> 11) * (1.0 / 9007199254740992.0);
        const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        out[e] = centre + sigma * (float)z;
    }
}

static int gen_grid(long long total) {
    long long g = (total + 255) / 256;
    return (int)(g > 65536 ? 65536 : (g < 1 ? 1 : g));
}

hipError_t launch_gen_uniform(float *out, int n, int dim, uint64_t seed, float scale,
                              hipStream_t s) {
    hipLaunchKernelGGL(gen_uniform_kernel, dim3(gen_grid((long long)n * dim)), dim3(256), 0, s,
                       out, n, dim, seed, scale);
    return hipGetLastError();
}

hipError_t launch_gen_mixture(float *out, int n, int dim, int ncomp, float sigma,
                              uint64_t centre_seed, uint64_t point_seed, hipStream_t s) {
    hipLaunchKernelGGL(gen_mixture_kernel, dim3(gen_grid((long long)n * dim)), dim3(256), 0, s,
                       out, n, dim, ncomp, sigma, centre_seed, point_seed);
    return hipGetLastError();
}

}  // namespace knnhip
