// Precision form of the block-index scan on block-centroid residuals:
// single-pass bfloat16 MFMA products (v_mfma_f32_32x32x16_bf16, FP32
// accumulation) of (q - c) and (r - c), c the centroid of the block a task
// scans.  This replaces the reference's FP32 contraction
// (code/knncuda.cu:744, cublasSgemm) for index searches with
// options.precision = 2.
//
// Key.  Per block b the index keeps c (FP32), every member's residual
// e_r = bf16(r - c) (rounded to nearest even, FP32 difference first) in the
// scan's tile layout, its FP32 squared norm n_r = fl(sum e_r^2) and the
// block radius rho = max ||r - c|| (exact residuals, FP64, rounded up).  A
// task forms e_q = bf16(q - c) for its queries, n_q = fl(sum e_q^2), and
//   key = n_r + n_q + sum (-2 e_q) e_r ,
// a FP32 evaluation of ||e_q - e_r||^2 (every bfloat16 product is exact in
// FP32).  With the residual roundings below delta = 2^-8 + 2^-23 of each
// coordinate, ||e_q - e_r|| is within eps = delta (||q - c|| + ||r - c||)
// of ||q - r||, and the key within eta = coef (||e_q|| + ||e_r||)^2 of
// ||e_q - e_r||^2 (coef: the FP32 norm and accumulation roundings, unit
// 2^-23 per MFMA term to allow for any rounding inside the MFMA).  Each task
// stores, with the heap row of every (query, block) pair it scans, eps, eta
// and L = ||q - c|| - rho, which bounds the distance of every member of the
// block (triangle inequality).  The re-rank's certificate then bounds the
// distance of every reference the scan rejected in that block by
// max(L, sqrt(tau - eta) - eps): far blocks, whose residuals are large, are
// covered by L.
//
// Residuals are small where blocks are tight (a cluster of the data), so the
// bound is far below that of bfloat16 products of the raw coordinates, whose
// error scales with ||q|| ||r||.
#include "knn_common.h"
#include "knn_heap.h"

namespace knnhip {

namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));

constexpr int kResMaxDim = 256;        // 16 chunks of B operands in VGPRs
constexpr float kResDelta = 0x1p-8f + 0x1p-23f;

__device__ __forceinline__ unsigned short bf_bits(__bf16 x) { return __builtin_bit_cast(unsigned short, x); }

// One workgroup per unit (a block, or a ball: a contiguous range of
// positions): its centroid (mean in FP64), the residual tiles and squared
// norms of its positions, and its radius.  Positions outside every unit keep
// the zeros the tiles were cleared to.
__global__ __launch_bounds__(256) void resid_build_kernel(const float *__restrict__ pm, int ld, int dim, int nch,
                                                          const int *__restrict__ block_start,
                                                          const int *__restrict__ block_count,
                                                          float *__restrict__ cent, unsigned short *__restrict__ tiles,
                                                          float *__restrict__ rnorm, float *__restrict__ rho) {
    __shared__ float c_s[kResMaxDim];
    __shared__ double red[256];
    const int b = blockIdx.x, t = threadIdx.x;
    const long long p0 = block_start[b];
    const int n = block_count[b];
    for (int d = t; d < ld; d += blockDim.x) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += (double)pm[(p0 + j) * ld + d];
        const float c = (d < dim && n > 0) ? (float)(s / (double)n) : 0.0f;
        cent[(long long)b * ld + d] = c;
        if (d < kResMaxDim) c_s[d] = c;
    }
    __syncthreads();
    double emax = 0.0;
    for (int j = t; j < n; j += blockDim.x) {
        const long long p = p0 + j;
        const long long tile = p / kRefsPerTile;
        const int m = (int)(p % kRefsPerTile);
        const bool member = true;
        const float *row = pm + p * ld;
        float nr = 0.0f;
        double ex = 0.0;
        for (int cc = 0; cc < nch; ++cc)
            for (int h = 0; h < 2; ++h) {
                unsigned short v[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int d = 16 * cc + 8 * h + e;
                    float y = 0.0f;
                    if (member && d < dim) {
                        const float r = row[d];
                        y = r - c_s[d];
                        const double ed = (double)r - (double)c_s[d];
                        ex += ed * ed;
                    }
                    const __bf16 yb = (__bf16)y;
                    const float yf = (float)yb;
                    nr += yf * yf;
                    v[e] = bf_bits(yb);
                }
                uint4 pk;
                pk.x = v[0] | ((unsigned)v[1] << 16);
                pk.y = v[2] | ((unsigned)v[3] << 16);
                pk.z = v[4] | ((unsigned)v[5] << 16);
                pk.w = v[6] | ((unsigned)v[7] << 16);
                *reinterpret_cast<uint4 *>(tiles + ((tile * nch + cc) * kWave + m + 32 * h) * 8) = pk;
            }
        rnorm[p] = nr;
        emax = fmax(emax, ex);
    }
    red[t] = emax;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w) red[t] = fmax(red[t], red[t + w]);
        __syncthreads();
    }
    if (t == 0) {
        const double r = sqrt(red[0]) * (1.0 + 1e-12);
        float rf = (float)r;
        if ((double)rf < r) rf = nextafterf(rf, INFINITY);
        rho[b] = rf;
    }
}

// Task scan (block-index tasks of up to 32 queries over one block), the
// structure of the bfloat16 three-term task scan with one product per 16
// dimensions on the residual tiles.  NB: chunks in the register ring (loads
// run NB - 1 chunks ahead).  IND: every wave is a task of its own, with its
// own heaps and queue and no workgroup barrier.  MERGE (one-wave tasks of up
// to kMergeQueries queries, kp <= kMergeMaxKp): sorted merged candidate
// lists in place of the heaps and queue; rows start at the task's list
// offset.
template <int NCH, int NB, bool IND, bool MERGE>
__global__ __launch_bounds__(256, 2) void scan_res_task_kernel(ScanArgs a) {
    extern __shared__ __align__(16) unsigned char smem_all[];
    static_assert(NCH % NB == 0 && NB >= 2, "ring position must repeat every step");
    static_assert(IND || !MERGE, "merged lists are per wave");
    const int nwaves = blockDim.x >> 6;
    const int wave_id = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 31, h = lane >> 5;
    const int waves = IND ? 1 : nwaves, wave = IND ? 0 : wave_id;
    // (wave-uniform, and read as such: the task's fields then load as scalars and
    // the tiles' buffer resource sits in scalar registers, as buffer loads need)
    const int tix = __builtin_amdgcn_readfirstlane(IND ? blockIdx.x * nwaves + wave_id : blockIdx.x);
    if (IND && tix >= a.ntasks) return;  // (no workgroup barrier in this form)
    const int kp = a.kp, kps = heap_stride(kp);
    const size_t wave_lds = MERGE ? merge_lds_bytes(kp) : scan_heap_bytes(kp) + kQueueBytes;
    unsigned char *smem = smem_all + (IND ? (size_t)wave_id * wave_lds : 0);

    const ScanTask tk = a.tasks[tix];
    const int r_begin = tk.ref_begin, r_end = tk.ref_end;
    // (merged lists: the task list holds at most kMergeQueries queries a task)
    const int nq = MERGE ? min(tk.qcount, kMergeQueries) : tk.qcount;
    if (nq <= 0) return;  // empty task: uniform over the block (IND: over the wave)
    const bool qvalid = col < nq;
    const int q = a.qlist[tk.qoff + (qvalid ? col : 0)];
    const int task_row = MERGE ? tk.qoff : tix * kQueriesPerWave;
    const int tid = IND ? lane : threadIdx.x, tstride = IND ? kWave : blockDim.x;
    auto task_sync = [&]() {
        if constexpr (IND) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else {
            __syncthreads();
        }
    };

    float *hv = reinterpret_cast<float *>(smem);
    int *hi = reinterpret_cast<int *>(smem + (size_t)kQueriesPerWave * kps * sizeof(float));
    int *lock = hi + (size_t)kQueriesPerWave * kps;
    KeyRef *lists = reinterpret_cast<KeyRef *>(smem);
    KeyRef *cand = lists + (size_t)kMergeQueries * kp;
    int mcount = 0, mfill = 0;  // (merged lists) this lane's pending candidates, its column's fill
    float mthr = INFINITY;      // (merged lists) its column's kp-th key
    if constexpr (MERGE) {
        for (int e = lane; e < nq * kp; e += kWave) {
            lists[e].k = INFINITY;
            lists[e].r = -1;
        }
    } else {
        for (int e = tid; e < kQueriesPerWave * kps; e += tstride) {
            hv[e] = INFINITY;
            hi[e] = -1;
        }
        if (tid < kQueriesPerWave) lock[tid] = 0;
    }
    task_sync();
    float *myv = hv + col * kps;
    float theta = INFINITY;
    float published = INFINITY;
    unsigned *theta_q = a.theta + q;
    unsigned char *qbase = smem + scan_heap_bytes(kp) + (size_t)wave * kQueueBytes;
    float *qv = reinterpret_cast<float *>(qbase);
    int *qr = reinterpret_cast<int *>(qbase + kQueueCap * sizeof(float));
    int *qc = reinterpret_cast<int *>(qbase + kQueueCap * 2 * sizeof(float));
    int qcount = 0;  // wave-uniform
    const unsigned long long lanes_below = (1ull << lane) - 1ull;

    // B operands: -2 e_q, e_q = bf16(q - c) of the task's block, lane
    // (col, h) holding dimensions 16c + 8h .. +7 of chunk c; n_q over both
    // lane halves
    const int blk = tk.unit;  // the task's block (or ball): its centroid and radius
    const float *cen = a.res_cent + (long long)blk * a.query_pm_ld;
    const float *qrow = a.query_pm + (long long)q * a.query_pm_ld;
    bf16x8_t bq[NCH];
    float nqh = 0.0f;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const int d4 = 16 * c + 8 * h + 4 * p;
            const float4 x = row_piece(qrow, d4, a.query_pm_ld), m = row_piece(cen, d4, a.query_pm_ld);
            const float y[4] = {x.x - m.x, x.y - m.y, x.z - m.z, x.w - m.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const __bf16 yb = (__bf16)y[e];
                const float yf = (float)yb;
                nqh += yf * yf;
                bq[c][4 * p + e] = (__bf16)(-2.0f * yf);
            }
        }
    }
    const float nqv = nqh + __shfl_xor(nqh, 32);
    // Bounds of this (query, block) pair, stored with its heap row:
    //   ||e_q~|| within sqrt(n_q (1 -/+ g_D)), and ||q - c|| within
    //   ||e_q~|| (1 -/+ delta)^-1;
    //   L: every member r has ||q - r|| >= ||q - c|| - rho (triangle
    //      inequality), rounded down;
    //   eps, eta: the residual rounding and key evaluation bounds.
    if (h == 0 && qvalid) {
        const float rho = a.res_rho[blk];
        const float eqt = sqrtf(nqv * a.res_norm_up) * (1.0f + 0x1p-22f);
        const float eq = eqt / (1.0f - kResDelta) * (1.0f + 0x1p-22f);
        const float eq_lo = sqrtf(nqv * a.res_norm_dn) * (1.0f - 0x1p-22f) / (1.0f + kResDelta) * (1.0f - 0x1p-22f);
        const float L = fmaxf(eq_lo - rho, 0.0f) * (1.0f - 0x1p-20f);
        const float eps = kResDelta * (eq + rho) * (1.0f + 0x1p-20f);
        const float sr = eqt + rho * (1.0f + kResDelta);
        const float eta = a.res_eta_coef * sr * sr * (1.0f + 0x1p-20f);
        *reinterpret_cast<float4 *>(a.res_row + 4 * ((long long)task_row + col)) = make_float4(L, eps, eta, 0.0f);
    }

    // A operands through one buffer resource over the task's tiles: loads
    // past them return zeros.
    constexpr long long kTileBytes = (long long)NCH * kWave * 16;
    const int span_refs = (r_end - r_begin + kRefsPerStep - 1) / kRefsPerStep * kRefsPerStep;
    const long long valid_tiles = max(0, min(span_refs, a.r_pad - r_begin)) / kRefsPerTile;
    const __amdgpu_buffer_rsrc_t rs =
        uniform_buffer_rsrc(a.packed_res + (long long)(r_begin / kRefsPerTile) * (kTileBytes / 2),
                            (unsigned)(valid_tiles * kTileBytes));
    auto lane_offsets = [&](int r_start, unsigned *vo) {
        const long long t0 = (r_start - r_begin) / kRefsPerTile;
#pragma unroll
        for (int mt = 0; mt < kMTiles; ++mt)
            vo[mt] = (unsigned)min((t0 + mt) * kTileBytes + lane * 16, 0xFFFFFFF0ll);
    };
    u32x4_t ring[NB][kMTiles];
    auto issue = [&](u32x4_t(&buf)[kMTiles], int c, const unsigned *vo) {
#pragma unroll
        for (int mt = 0; mt < kMTiles; ++mt)
            buf[mt] = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)vo[mt], c * kWave * 16, 0);
    };

    const int wstride = waves * kRefsPerStep;
    const int r_first = r_begin + wave * kRefsPerStep;
    unsigned voff[kMTiles];
    lane_offsets(r_first, voff);
    if (r_first < r_end) {
#pragma unroll
        for (int j = 0; j < NB - 1; ++j) issue(ring[j], j, voff);
    }

    for (int r0 = r_first; r0 < r_end; r0 += wstride) {
        f32x16_t acc[kMTiles];
        // start from n_r + n_q: element i of tile mt is position
        // r0 + 32 mt + 8 (i / 4) + 4h + i % 4
        if (r0 + kRefsPerStep <= a.r_pad) {
#pragma unroll
            for (int mt = 0; mt < kMTiles; ++mt)
#pragma unroll
                for (int g4 = 0; g4 < 4; ++g4) {
                    const float4 nr = *reinterpret_cast<const float4 *>(
                        a.res_norm + r0 + kRefsPerTile * mt + 8 * g4 + 4 * h);
                    acc[mt][4 * g4 + 0] = nr.x + nqv;
                    acc[mt][4 * g4 + 1] = nr.y + nqv;
                    acc[mt][4 * g4 + 2] = nr.z + nqv;
                    acc[mt][4 * g4 + 3] = nr.w + nqv;
                }
        } else {
#pragma unroll
            for (int mt = 0; mt < kMTiles; ++mt)
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int r = r0 + kRefsPerTile * mt + acc_row(i, h);
                    acc[mt][i] = r < a.r_pad ? a.res_norm[r] + nqv : INFINITY;
                }
        }
        const unsigned tq = __hip_atomic_load(theta_q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned voff_next[kMTiles];
        lane_offsets(r0 + wstride, voff_next);
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const int ci = c + NB - 1;
            if (ci < NCH)
                issue(ring[ci % NB], ci, voff);
            else
                issue(ring[ci % NB], ci - NCH, voff_next);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int mt = 0; mt < kMTiles; ++mt)
                acc[mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, ring[c % NB][mt]),
                                                                  bq[c], acc[mt], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int mt = 0; mt < kMTiles; ++mt) voff[mt] = voff_next[mt];

        // Epilogue (as the three-term task scan): keys below the freshest
        // threshold go to the wave's queue, drained into the heaps (merged
        // lists: to the lane's candidate slots, merged every step).
        theta = fminf(fminf(theta, ordered_to_float(tq)), MERGE ? mthr : shared_root(myv));
        if (r0 + kRefsPerStep > r_end || r0 < tk.ref_lo) {
            // rows outside the task take no part below
#pragma unroll
            for (int mt = 0; mt < kMTiles; ++mt)
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int r = r0 + kRefsPerTile * mt + acc_row(i, h);
                    if (r >= r_end || r < tk.ref_lo) acc[mt][i] = INFINITY;
                }
        }
        float kmin = INFINITY;
#pragma unroll
        for (int mt = 0; mt < kMTiles; ++mt)
#pragma unroll
            for (int i = 0; i < 16; ++i) kmin = fminf(kmin, acc[mt][i]);
        if (MERGE && __ballot(qvalid && kmin < theta)) {
            merge_step(acc, theta, kp, qvalid, r0, h, col, lane, nq, lists, cand, mcount, mfill, mthr);
            if (h == 0 && qvalid && theta < published) {
                published = theta;
                __hip_atomic_fetch_min(theta_q, float_to_ordered(theta), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
        } else if (!MERGE && __ballot(qvalid && kmin < theta)) {
            int ncand = 0;
#pragma unroll
            for (int mt = 0; mt < kMTiles; ++mt)
#pragma unroll
                for (int i = 0; i < 16; ++i) ncand += acc[mt][i] < theta ? 1 : 0;
            if (!qvalid) ncand = 0;
            const int npair = ncand + __shfl_xor(ncand, 32);
            float cut = theta;
            if (__ballot(npair >= 2 * kp)) cut = step_cut(acc, theta, kp, npair >= 2 * kp);
#pragma unroll
            for (int mt = 0; mt < kMTiles; ++mt) {
                bool cm = false;
#pragma unroll
                for (int i = 0; i < 16; ++i) cm |= acc[mt][i] < cut;
                if (__ballot(cm && qvalid) == 0) continue;
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    if ((i & 7) == 0 && qcount > kQueueCap - 8 * kWave) {
                        drain_queue(qv, qr, qc, qcount, hv, hi, lock, kps, kp, lane);
                        qcount = 0;
                        theta = fminf(theta, shared_root(myv));
                    }
                    const int r = r0 + kRefsPerTile * mt + acc_row(i, h);
                    const bool take = qvalid && acc[mt][i] < cut && r < r_end;
                    const unsigned long long m = __ballot(take);
                    if (m == 0) continue;
                    if (take) {
                        const int pos = qcount + __popcll(m & lanes_below);
                        qv[pos] = acc[mt][i];
                        qr[pos] = r;
                        qc[pos] = col;
                    }
                    qcount += __popcll(m);
                }
            }
            // a heap still short of kp entries (root +inf) lets every key of
            // the next step through: fill it now
            const bool young = __ballot(qvalid && theta == INFINITY) != 0;
            if (qcount >= kQueueCap / 2 || (young && qcount > 0)) {
                drain_queue(qv, qr, qc, qcount, hv, hi, lock, kps, kp, lane);
                qcount = 0;
                theta = fminf(theta, shared_root(myv));
                if (h == 0 && qvalid && theta < published) {
                    published = theta;
                    __hip_atomic_fetch_min(theta_q, float_to_ordered(theta), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
    }
    if constexpr (MERGE) {
        // (every step with candidates merged them; the lists are final)
        task_sync();
        const long long obase = (long long)task_row * kp;
        for (int e = lane; e < nq * kp; e += kWave) {
            const KeyRef v = lists[e];
            a.part_val[obase + e] = v.k;
            a.part_idx[obase + e] = v.r;
        }
        return;
    }
    drain_queue(qv, qr, qc, qcount, hv, hi, lock, kps, kp, lane);
    theta = fminf(theta, shared_root(myv));
    if (h == 0 && qvalid && theta < published)
        __hip_atomic_fetch_min(theta_q, float_to_ordered(theta), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

    task_sync();
    const long long obase = (long long)task_row * kp;
    for (int e = tid; e < nq * kp; e += tstride) {
        const int row = e / kp, j = e - row * kp;
        a.part_val[obase + e] = hv[row * kps + j];
        a.part_idx[obase + e] = hi[row * kps + j];
    }
}

}  // namespace

#define KNN_RES_TASK_INSTANCES(X) \
    X(2, 2)                       \
    X(4, 2)                       \
    X(6, 3)                       \
    X(8, 4)                       \
    X(12, 4)                      \
    X(16, 2)

int scan_res_nch(int dim) {
    const int n = (dim + 15) / 16;
    for (int c : {2, 4, 6, 8, 12, 16})
        if (c >= n) return c;
    return 0;  // wider dimensions: the B operands would not fit in VGPRs
}

// Key error coefficient of the residual scan over nch chunks: the FP32
// norms of e_q and e_r (dim terms) and the accumulation of 16 nch products
// plus the starting value (unit 2^-23), relative to (||e_q|| + ||e_r||)^2,
// widened by 5 %.
float res_eta_coeff(int dim, int nch) {
    const double n = 16.0 * nch + 2.0;
    const double gn = n * 0x1p-23 / (1.0 - n * 0x1p-23);
    const double gd = dim * 0x1p-24 / (1.0 - dim * 0x1p-24);
    return (float)(1.05 * (gn + gd + 0x1p-24)) * (1.0f + 0x1p-20f);
}

// Multiplier of an FP32 sum of dim non-negative terms that bounds the exact
// sum from above: 1 / (1 - g_dim), rounded up.
float res_norm_up(int dim) {
    const double gd = dim * 0x1p-24 / (1.0 - dim * 0x1p-24);
    return (float)(1.0 / (1.0 - gd)) * (1.0f + 0x1p-22f);
}

// ... and from below: 1 / (1 + g_dim), rounded down.
float res_norm_dn(int dim) {
    const double gd = dim * 0x1p-24 / (1.0 - dim * 0x1p-24);
    return (float)(1.0 / (1.0 + gd)) * (1.0f - 0x1p-22f);
}

hipError_t launch_resid_build(const float *pm, int ld, int dim, int nch, const int *unit_start,
                              const int *unit_count, int nunits, float *cent, unsigned short *tiles, float *rnorm,
                              float *rho, hipStream_t s) {
    if (dim > kResMaxDim || 16 * nch < dim || ld % 4) return hipErrorInvalidValue;
    hipLaunchKernelGGL(resid_build_kernel, dim3(nunits), dim3(256), 0, s, pm, ld, dim, nch, unit_start, unit_count,
                       cent, tiles, rnorm, rho);
    return hipGetLastError();
}

hipError_t launch_scan_res_task(const ScanArgs &a, int nch, int waves, size_t lds, hipStream_t s,
                                int tasks_per_group, bool merge) {
    if (!a.tasks || !a.packed_res || !a.res_norm || !a.res_cent || !a.res_rho || !a.res_row || !a.query_pm)
        return hipErrorInvalidValue;
    if (tasks_per_group > 1 && (waves != 1 || tasks_per_group > 4)) return hipErrorInvalidValue;
    if (merge && (waves != 1 || tasks_per_group < 1 || a.kp > kMergeMaxKp)) return hipErrorInvalidValue;
    const size_t lds_ind = (size_t)tasks_per_group * (merge ? merge_lds_bytes(a.kp) : scan_heap_bytes(a.kp) + kQueueBytes);
    const dim3 grid_ind((a.ntasks + tasks_per_group - 1) / tasks_per_group), block_ind(tasks_per_group * kWave);
#define LAUNCH_IND(N, B, M)                                                                        \
    {                                                                                              \
        const void *fi = (const void *)scan_res_task_kernel<N, B, true, M>;                        \
        if (lds_ind > 65536) {                                                                     \
            hipError_t e = hipFuncSetAttribute(fi, hipFuncAttributeMaxDynamicSharedMemorySize,      \
                                               (int)lds_ind);                                      \
            if (e != hipSuccess) return e;                                                         \
        }                                                                                          \
        hipLaunchKernelGGL((scan_res_task_kernel<N, B, true, M>), grid_ind, block_ind, lds_ind, s, a); \
        return hipGetLastError();                                                                  \
    }
#define X(N, B)                                                                                    \
    if (nch == N) {                                                                                \
        if (merge) LAUNCH_IND(N, B, true)                                                          \
        if (tasks_per_group > 1) LAUNCH_IND(N, B, false)                                           \
        const void *fn = (const void *)scan_res_task_kernel<N, B, false, false>;                   \
        if (lds > 65536) {                                                                         \
            hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,      \
                                               (int)lds);                                          \
            if (e != hipSuccess) return e;                                                         \
        }                                                                                          \
        hipLaunchKernelGGL((scan_res_task_kernel<N, B, false, false>), dim3(a.ntasks), dim3(waves * kWave), lds, \
                           s, a);                                                                  \
        return hipGetLastError();                                                                  \
    }
    KNN_RES_TASK_INSTANCES(X)
#undef X
#undef LAUNCH_IND
    return hipErrorInvalidValue;
}

}  // namespace knnhip
