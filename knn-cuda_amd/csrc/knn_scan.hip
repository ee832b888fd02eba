// FP32 MFMA distance scan fused with a streaming per-query top-kp selection.
//
// Replaces the reference's distance + sort kernels (code/knncuda.cu:20-91
// compute_distances, :105-122 compute_distance_texture, :143-189
// modified_insertion_sort, :217-250 squared norms / cuBLAS form): instead of
// materialising the ref_nb x query_nb matrix, every wave keeps its 32 queries'
// best kp approximate keys in LDS max-heaps while it streams reference tiles.
//
// Contraction: for reference r and query q the wave computes on MFMA
//     key(r, q) = ||r||^2 - 2 r.q          (= ||r - q||^2 - ||q||^2)
// with v_mfma_f32_32x32x2_f32.  A = references from the packed tile layout
// (below), B = -2 q resident in VGPRs for the whole scan, and one extra
// k-step that adds ||r||^2 (A = norm, B = 1).  The per-query constant
// ||q||^2 does not change the order and is added back by the re-rank stage.
//
// Packed reference layout (written by the norms pass): for tile t of 32
// points and dimension octet j, 64 lanes x 4 floats, lane m + 32h holding
// point 32t+m, dimensions 8j+4h .. 8j+4h+3 (zero past dim and past R).  One
// 16-byte load per lane feeds four consecutive k-steps: k-step 4j+u reads
// dimension 8j + 4h + u on lane half h, and the B operand uses the same map.
//
// Selection: the accumulator holds 16 references of one query per lane.  A
// key below the query's threshold theta is a candidate; the common case is
// one compare per key and a wave ballot.  Candidates replace the root of the
// query's max-heap (kp entries, LDS) and sift down.  theta is the smaller of
// the heap root and the query's shared threshold, the minimum root published
// by every slice scanning that query (an agent-scope atomic min), so slices
// prune with each other's progress.  Invariant used by the certificate:
// every reference left out of the heaps has key >= the final shared
// threshold, or >= the kp-th smallest key kept.
#include "knn_common.h"
#include "knn_heap.h"

namespace knnhip {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// KS: k-steps per dimension chunk (B operands held in VGPRs), a multiple of 4.
// NBUF: ring of one-octet (4 k-step) load buffers; loads run NBUF-1 octets
// ahead of the MFMAs that consume them.
//
// Block = W waves on the same 32 queries; wave w takes the slice's reference
// steps w, w+W, w+2W, ...  All W waves feed one heap per query in LDS,
// guarded by a per-query LDS lock, so every query's threshold is the kp-th
// smallest key over everything the block has scanned so far.
// BMODE: where the B operands live (kBRegs, kBChunks, kBLds).
// MAXW: most waves per block (4, or 8 where the heaps of large kp leave LDS
// for one block per CU: eight waves then keep two per SIMD).
// scan_f32_block: one block's task t (of a.tasks, or of the slice x query
// block grid); scan_f32_kernel runs task blockIdx.x, the batched kernel
// below the tasks it takes from a work queue.  smem: the block's dynamic LDS.
template <int KS, int NBUF, int BMODE, int MAXW>
__device__ __forceinline__ void scan_f32_block(const ScanArgs &a, const int t, unsigned char *smem) {
    constexpr int OCT = KS / 4;  // octets per chunk
    constexpr bool MULTI = BMODE == kBChunks;
    constexpr bool BLDS = BMODE == kBLds;
    constexpr int QCAP = scan_queue_cap(MAXW);
    constexpr size_t QBYTES = scan_queue_bytes(MAXW);
    static_assert(KS % 4 == 0, "k-steps come in octets of 4");
    static_assert(OCT % NBUF == 0, "ring position must repeat every chunk");

    const int waves = blockDim.x >> 6;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int col = lane & 31;
    const int h = lane >> 5;

    // One task per block: a reference range and up to 32 queries; its heaps
    // leave as output rows task_row .. task_row+31.
    int r_begin, r_end, nq, task_row;
    int r_lo = 0;  // rows below r_lo (a task's tile-rounded start) take no part
    const int *qids = nullptr;  // query ids of the task (nullptr: consecutive from qfirst)
    int qfirst = 0;
    if (a.tasks) {
        const ScanTask tk = a.tasks[t];
        r_begin = tk.ref_begin;
        r_end = tk.ref_end;
        r_lo = tk.ref_lo;
        nq = tk.qcount;
        qids = a.qlist + tk.qoff;
        task_row = t * kQueriesPerWave;
    } else {
        int slice, qb;
        if (a.xcd_map) {
            // Consecutive block ids land on different XCDs; give every block of
            // one slice the same XCD so the slice is read from HBM once per L2.
            const int xcd = t & 7, j = t >> 3;
            qb = j % a.nqb;
            slice = a.slice0 + (j / a.nqb) * 8 + xcd;
        } else {
            slice = a.slice0 + t / a.nqb;
            qb = t % a.nqb;
        }
        qfirst = qb * kQueriesPerWave;
        nq = min(kQueriesPerWave, a.Q - qfirst);
        r_begin = slice * a.slice_len;
        r_end = min(a.R, r_begin + a.slice_len);
        task_row = (slice * a.nqb + qb) * kQueriesPerWave;
    }
    if (nq <= 0) return;  // whole block is past the last query
    // Columns past the task's queries compute query 0's keys (their heaps are
    // never written out) and take no candidates.
    const bool qvalid = col < nq;
    const int qsel = qvalid ? col : 0;
    const int q = qids ? qids[qsel] : qfirst + qsel;

    // Shared heaps (odd row stride: the lanes of a half walk 32 heaps in
    // step on distinct banks) and one lock word per query.
    const int kp = a.kp;
    const int kps = heap_stride(kp);
    float *hv = reinterpret_cast<float *>(smem);
    int *hi = reinterpret_cast<int *>(smem + (size_t)kQueriesPerWave * kps * sizeof(float));
    int *lock = hi + (size_t)kQueriesPerWave * kps;
    for (int e = threadIdx.x; e < kQueriesPerWave * kps; e += blockDim.x) {
        hv[e] = INFINITY;
        hi[e] = -1;
    }
    if (threadIdx.x < kQueriesPerWave) lock[threadIdx.x] = 0;
    // kBLds: the block's 32 query rows (-2q, zero past dim and past the
    // task's queries) after the heaps and queues, row pitch 8 oct + 4 floats
    float *bl = nullptr;
    const int bpitch = 8 * a.oct + 4;
    if constexpr (BLDS) {
        bl = reinterpret_cast<float *>(smem + scan_heap_bytes(kp) + (size_t)waves * QBYTES);
        const int dpad = 8 * a.oct;
        if (a.query_pm) {
            // point-major rows: consecutive threads take consecutive dims of a row
            for (int e = threadIdx.x; e < kQueriesPerWave * dpad; e += blockDim.x) {
                const int qi = e / dpad, d = e - qi * dpad;
                float v = 0.0f;
                if (qi < nq && d < a.dim)
                    v = -2.0f * a.query_pm[(long long)(qids ? qids[qi] : qfirst + qi) * a.query_pm_ld + d];
                bl[qi * bpitch + d] = v;
            }
        } else {
            // dim-major columns: consecutive threads take consecutive queries of a dim
            for (int e = threadIdx.x; e < kQueriesPerWave * dpad; e += blockDim.x) {
                const int d = e / kQueriesPerWave, qi = e - d * kQueriesPerWave;
                float v = 0.0f;
                if (qi < nq && d < a.dim)
                    v = -2.0f * a.query[(long long)d * a.ldq + (qids ? qids[qi] : qfirst + qi)];
                bl[qi * bpitch + d] = v;
            }
        }
    }
    __syncthreads();
    float *myv = hv + col * kps;
    float theta = INFINITY;      // threshold in use (<= the shared heap root)
    float published = INFINITY;  // last value pushed to the cross-slice threshold
    unsigned *theta_q = a.theta + q;

    // Per-wave candidate queue (key, reference, query column) in LDS, after
    // the heaps and locks.
    unsigned char *qbase_ptr = smem + scan_heap_bytes(kp) + (size_t)wave * QBYTES;
    float *qv = reinterpret_cast<float *>(qbase_ptr);
    int *qr = reinterpret_cast<int *>(qbase_ptr + QCAP * sizeof(float));
    int *qc = reinterpret_cast<int *>(qbase_ptr + QCAP * 2 * sizeof(float));
    int qcount = 0;  // wave-uniform
    const unsigned long long lanes_below = (1ull << lane) - 1ull;
    // Insert every queued candidate still below its query's heap root.  The
    // wave takes 64 entries at a time, one per lane; entries of one query are
    // serialised by the query's lock.  The retry loop is wave-uniform (its
    // condition is a ballot) and each lane makes one attempt per iteration,
    // releasing a lock in the iteration that took it: a per-lane spin loop
    // would let the compiler move a winner's critical section past the loop
    // while a losing lane of the same wave spins on that lock.
    auto drain = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        for (int base = 0; base < qcount; base += kWave) {
            const int j = base + lane;
            bool pending = j < qcount;
            const float v = pending ? qv[j] : 0.0f;
            const int r = pending ? qr[j] : 0;
            const int c = pending ? qc[j] : 0;
            float *cv = hv + c * kps;
            int *ci = hi + c * kps;
            while (__ballot(pending)) {
                if (pending) {
                    if (!(v < shared_root(cv))) {
                        pending = false;
                    } else if (try_lock(lock + c)) {
                        if (v < cv[0]) heap_replace_top(cv, ci, kp, v, r);
                        unlock(lock + c);
                        pending = false;
                    }
                }
            }
        }
        qcount = 0;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };

    const long long ldq = a.ldq;
    const float *__restrict__ qcol = a.query + q;
    float bq[KS];
    auto load_b = [&](int c) {
        if (a.query_pm) {
            // k-steps 4j .. 4j+3 read dims 8j + 4h .. +3 of the query: one
            // 16-byte piece of its row (whole pieces: the row stride is a
            // multiple of 4 and the pads are zero).  The chunked instance
            // reloads these every chunk: KS / 4 row pieces instead of KS
            // strided column reads.
            const float *row = a.query_pm + (long long)q * a.query_pm_ld;
#pragma unroll
            for (int j = 0; j < KS / 4; ++j) {
                const int d4 = c * 2 * KS + 8 * j + 4 * h;
                const float4 v = d4 < a.query_pm_ld ? *reinterpret_cast<const float4 *>(row + d4)
                                                    : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                bq[4 * j + 0] = -2.0f * v.x;
                bq[4 * j + 1] = -2.0f * v.y;
                bq[4 * j + 2] = -2.0f * v.z;
                bq[4 * j + 3] = -2.0f * v.w;
            }
            return;
        }
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) {
            // Branch-free: clamp the row, then zero the padding dimensions.
            const int d = c * 2 * KS + 8 * (kk >> 2) + 4 * h + (kk & 3);
            const float v = qcol[(long long)min(d, a.dim - 1) * ldq];
            bq[kk] = d < a.dim ? -2.0f * v : 0.0f;
        }
    };
    if (BMODE == kBRegs) load_b(0);
    const int nchunks = MULTI ? a.nchunks : 1;

    const int wstride = waves * kRefsPerStep;  // a wave's distance between steps

    // Packed operands through one buffer resource per slice; the lane's
    // offset is a VGPR, the octet offset a scalar.  The resource covers only
    // the slice's own tiles, so the prefetch past a wave's last step (issued
    // unconditionally, to keep the load counter static) falls outside it and
    // returns zeros without touching memory.
    const long long tile_bytes = (long long)a.oct * kWave * 16;
    const int span_refs = (r_end - r_begin + kRefsPerStep - 1) / kRefsPerStep * kRefsPerStep;
    const long long valid_tiles = max(0, min(span_refs, a.r_pad - r_begin)) / kRefsPerTile;
    const __amdgpu_buffer_rsrc_t rs =
        uniform_buffer_rsrc(a.packed + (long long)(r_begin / kRefsPerTile) * a.oct * kWave * 4,
                            (unsigned)(valid_tiles * tile_bytes));
    auto lane_offsets = [&](int r_start, unsigned *vo) {
        const long long t0 = (r_start - r_begin) / kRefsPerTile;
#pragma unroll
        for (int mt = 0; mt < kMTiles; ++mt)
            vo[mt] = (unsigned)min((t0 + mt) * tile_bytes + lane * 16, 0xFFFFFFF0ll);
    };
    u32x4 abuf[NBUF][kMTiles];
    auto issue = [&](u32x4(&buf)[kMTiles], int oct, const unsigned *vo) {
#pragma unroll
        for (int mt = 0; mt < kMTiles; ++mt)
            buf[mt] = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)vo[mt], oct * kWave * 16, 0);
    };

    const int r_first = r_begin + wave * kRefsPerStep;
    unsigned voff[kMTiles];
    lane_offsets(r_first, voff);
    if (!MULTI && r_first < r_end) {
#pragma unroll
        for (int o = 0; o < NBUF - 1; ++o) issue(abuf[o], o, voff);
    }

    for (int r0 = r_first; r0 < r_end; r0 += wstride) {
        f32x16 acc[kMTiles];
        float nrm[kMTiles];
        const unsigned tq = __hip_atomic_load(theta_q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int mt = 0; mt < kMTiles; ++mt) {
            acc[mt] = f32x16{};
            const int r = r0 + kRefsPerTile * mt + col;
            nrm[mt] = r < a.R ? a.rnorm[r] : INFINITY;
        }
        const int r_next = r0 + wstride;
        unsigned voff_next[kMTiles];
        lane_offsets(r_next, voff_next);
        if constexpr (BLDS) {
            // every octet of the padded dimension (a multiple of NBUF octets) in
            // one pass; octet o sits in ring slot o % NBUF, and its B operands
            // are read from LDS one octet ahead
            const int octs = a.oct;
            float4 bcur = *reinterpret_cast<const float4 *>(bl + col * bpitch + 4 * h);
            for (int o0 = 0; o0 < octs; o0 += NBUF) {
#pragma unroll
                for (int j = 0; j < NBUF; ++j) {
                    const int o = o0 + j;
                    const int oi = o + NBUF - 1;
                    if (oi < octs)
                        issue(abuf[(j + NBUF - 1) % NBUF], oi, voff);
                    else
                        issue(abuf[(j + NBUF - 1) % NBUF], oi - octs, voff_next);
                    const int on = o + 1 < octs ? o + 1 : 0;
                    const float4 bnext =
                        *reinterpret_cast<const float4 *>(bl + col * bpitch + 8 * on + 4 * h);
                    __builtin_amdgcn_sched_barrier(0);
                    const float bv[4] = {bcur.x, bcur.y, bcur.z, bcur.w};
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
#pragma unroll
                        for (int mt = 0; mt < kMTiles; ++mt)
                            acc[mt] = __builtin_amdgcn_mfma_f32_32x32x2f32(
                                __uint_as_float(abuf[j][mt][u]), bv[u], acc[mt], 0, 0, 0);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                    bcur = bnext;
                }
            }
        }
        for (int c = 0; c < (BLDS ? 0 : nchunks); ++c) {
            if (MULTI) {
                load_b(c);
#pragma unroll
                for (int o = 0; o < NBUF - 1; ++o) issue(abuf[o], c * OCT + o, voff);
            }
#pragma unroll
            for (int o = 0; o < OCT; ++o) {
                constexpr int ahead = NBUF - 1;
                const int oi = o + ahead;
                if (oi < OCT)
                    issue(abuf[oi % NBUF], c * OCT + oi, voff);
                else if (!MULTI)
                    issue(abuf[oi % NBUF], oi - OCT, voff_next);
                __builtin_amdgcn_sched_barrier(0);
                // One k-step across the four tiles per group: consecutive
                // MFMAs always target different accumulators.
#pragma unroll
                for (int u = 0; u < 4; ++u) {
#pragma unroll
                    for (int mt = 0; mt < kMTiles; ++mt)
                        acc[mt] = __builtin_amdgcn_mfma_f32_32x32x2f32(
                            __uint_as_float(abuf[o % NBUF][mt][u]), bq[4 * o + u], acc[mt], 0, 0,
                            0);
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        }
        // ||r||^2 enters as one more k-step: A = norm (lane half 0), B = 1.
#pragma unroll
        for (int mt = 0; mt < kMTiles; ++mt)
            acc[mt] = __builtin_amdgcn_mfma_f32_32x32x2f32(h ? 0.0f : nrm[mt], h ? 0.0f : 1.0f,
                                                           acc[mt], 0, 0, 0);
#pragma unroll
        for (int mt = 0; mt < kMTiles; ++mt) voff[mt] = voff_next[mt];

        // Epilogue: one compare per key against the freshest threshold (the
        // shared heap root and the cross-slice value).  Keys below it are
        // appended to the wave's candidate queue; the queue is drained into
        // the heaps by all lanes at once when it fills.
        theta = fminf(fminf(theta, ordered_to_float(tq)), shared_root(myv));
        if (r0 + kRefsPerStep > r_end || r0 < r_lo) {
            // rows outside the task take no part below
#pragma unroll
            for (int mt = 0; mt < kMTiles; ++mt)
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int r = r0 + kRefsPerTile * mt + acc_row(i, h);
                    if (r >= r_end || r < r_lo) acc[mt][i] = INFINITY;
                }
        }
        // fast path: the step's smallest key against theta (a min3 chain);
        // the candidates are counted only on steps that have one
        float kmin = INFINITY;
#pragma unroll
        for (int mt = 0; mt < kMTiles; ++mt)
#pragma unroll
            for (int i = 0; i < 16; ++i) kmin = fminf(kmin, acc[mt][i]);
        if (__ballot(qvalid && kmin < theta)) {
            int ncand = 0;
#pragma unroll
            for (int mt = 0; mt < kMTiles; ++mt)
#pragma unroll
                for (int i = 0; i < 16; ++i) ncand += acc[mt][i] < theta ? 1 : 0;
            if (!qvalid) ncand = 0;
            // many candidates (a young heap): queue only a step-local cut
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
                for (int g = 0; g < 16 / kQueueGroup; ++g) {
                    if (__ballot(qvalid && group_min(acc[mt], g) < cut) == 0) continue;
                    if (qcount > QCAP - kQueueGroup * kWave) {
                        drain();
                        theta = fminf(theta, shared_root(myv));
                    }
#pragma unroll
                    for (int e = 0; e < kQueueGroup; ++e) {
                        const int i = kQueueGroup * g + e;
                        const int r = r0 + kRefsPerTile * mt + acc_row(i, h);
                        const bool take = qvalid && acc[mt][i] < cut && r < r_end;
                        const unsigned long long m = __ballot(take);
                        if (m == 0) continue;
                        const int n = __popcll(m);
                        if (take) {
                            const int pos = qcount + __popcll(m & lanes_below);
                            qv[pos] = acc[mt][i];
                            qr[pos] = r;
                            qc[pos] = col;
                        }
                        qcount += n;
                    }
                }
            }
            // Drain once a wave's pass of keys is queued: the heap root that
            // filters the next steps' keys (and the threshold published to
            // other slices' blocks) falls only at a drain, so a fuller queue
            // filters with a stale root (64 measured best against 16, 32 and
            // half the queue).
            if (qcount >= kWave) {
                drain();
                theta = fminf(theta, shared_root(myv));
                if (h == 0 && qvalid && theta < published) {
                    published = theta;
                    __hip_atomic_fetch_min(theta_q, float_to_ordered(theta), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
    }
    drain();
    theta = fminf(theta, shared_root(myv));
    if (h == 0 && qvalid && theta < published)
        __hip_atomic_fetch_min(theta_q, float_to_ordered(theta), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);

    // All waves done: the task's heaps are rows task_row .. task_row+nq-1
    // of the [row][kp] output.
    __syncthreads();
    const long long obase = (long long)task_row * kp;
    const int nvalid = nq * kp;
    for (int e = threadIdx.x; e < nvalid; e += blockDim.x) {
        const int row = e / kp, j = e - row * kp;
        a.part_val[obase + e] = hv[row * kps + j];
        a.part_idx[obase + e] = hi[row * kps + j];
    }
}

template <int KS, int NBUF, int BMODE, int MAXW>
__global__ __launch_bounds__(MAXW * 64, 8 / MAXW) void scan_f32_kernel(ScanArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    scan_f32_block<KS, NBUF, BMODE, MAXW>(a, blockIdx.x, smem);
}

// Batched scan: one persistent launch over the tasks of many problem
// classes.  Class c's launch arguments are cls[c] (its own packed tiles,
// queries, thresholds, task list and heap rows, its kp) and its instance
// variant[c] (kMultiVariants: the padded dimension's k-steps in VGPRs for
// 16, 32, 64 or 128 dimensions, else 128-dimension chunks).  A block takes
// work items (class, task of that class) from the queue `next` until it is
// empty; every block runs kMultiWaves waves whatever the class, with the
// short candidate queues (scan_queue_cap(8)) so that the heaps of a large
// kp leave two blocks per CU.  Larger items are queued first.
constexpr int kMultiWaves = 4;
__global__ __launch_bounds__(kMultiWaves * 64, 2) void scan_f32_multi_kernel(const ScanArgs *__restrict__ cls,
                                                                             const int *__restrict__ variant,
                                                                             const int2 *__restrict__ work,
                                                                             int nwork, unsigned *__restrict__ next) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int item;
    while (true) {
        __syncthreads();  // the previous item's LDS (heaps, queues, item) is free
        if (threadIdx.x == 0) item = (int)atomicAdd(next, 1u);
        __syncthreads();
        // (uniform, read as such: the item's arguments load as scalars and its
        // tiles' buffer resource sits in scalar registers, as buffer loads need)
        const int it = __builtin_amdgcn_readfirstlane(item);
        if (it >= nwork) break;  // uniform over the block
        const int2 w = work[it];
        const ScanArgs &a = cls[w.x];  // (read in place: uniform loads, no private copy)
        switch (variant[w.x]) {
            case 0: scan_f32_block<8, 2, kBRegs, 8>(a, w.y, smem); break;
            case 1: scan_f32_block<16, 4, kBRegs, 8>(a, w.y, smem); break;
            case 2: scan_f32_block<32, 4, kBRegs, 8>(a, w.y, smem); break;
            case 3: scan_f32_block<64, 4, kBRegs, 8>(a, w.y, smem); break;
            default: scan_f32_block<64, 2, kBChunks, 8>(a, w.y, smem); break;
        }
    }
}

int scan_multi_variant(int dim, int *ks, int *nchunks) {
    static const int kRegsKS[] = {8, 16, 32, 64};
    for (int v = 0; v < 4; ++v)
        if (2 * kRegsKS[v] >= dim) {
            *ks = kRegsKS[v];
            *nchunks = 1;
            return v;
        }
    *ks = 64;
    *nchunks = (dim + 127) / 128;
    return 4;
}

size_t scan_multi_lds(int kp) { return scan_heap_bytes(kp) + (size_t)kMultiWaves * scan_queue_bytes(8); }

int scan_multi_blocks_per_cu(size_t lds) {
    const void *fn = (const void *)scan_f32_multi_kernel;
    if (lds > 65536 && hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return 0;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, kMultiWaves * kWave, lds) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

hipError_t launch_scan_multi(const ScanArgs *cls, const int *variant, const int2 *work, int nwork, unsigned *next,
                             size_t lds, int grid, hipStream_t s) {
    if (grid < 1 || nwork < 1) return hipErrorInvalidValue;
    const void *fn = (const void *)scan_f32_multi_kernel;
    if (lds > 65536) {
        const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(scan_f32_multi_kernel, dim3(grid), dim3(kMultiWaves * kWave), lds, s, cls, variant, work,
                       nwork, next);
    return hipGetLastError();
}

// Dimension k-steps with an instantiated single-chunk scan (padded dim = 2*KS).
static const int kScanKS[] = {8, 16, 24, 32, 48, 64, 96};

int scan_pick_ks(int dim) {
    for (int ks : kScanKS)
        if (2 * ks >= dim) return ks;
    return 0;  // dimension needs the chunked scan
}

size_t scan_lds_bytes(int kp, int waves) {
    return scan_heap_bytes(kp) + (size_t)waves * scan_queue_bytes(waves);
}

// The same, with the staged query rows of the kBLds instance (oct octets).
size_t scan_block_lds(int kp, int waves, int bmode, int oct) {
    return scan_lds_bytes(kp, waves) +
           (bmode == kBLds ? (size_t)kQueriesPerWave * (8 * oct + 4) * sizeof(float) : 0);
}

// Ring depth per instance: 4 octets when the chunk's octet count allows it.
// The kBLds instance's KS only sets its ring (any multiple of 4 octets).
#define KNN_SCAN_INSTANCES(X) \
    X(8, 2, kBRegs)           \
    X(16, 4, kBRegs)          \
    X(24, 2, kBRegs)          \
    X(32, 4, kBRegs)          \
    X(48, 4, kBRegs)          \
    X(64, 4, kBRegs)          \
    X(96, 2, kBRegs)          \
    X(64, 2, kBChunks)        \
    X(16, 4, kBLds)

static const void *scan_fn(int ks, int bmode, int waves) {
#define X(K, N, M)                                                                         \
    if (ks == K && bmode == M)                                                             \
        return waves > 4 ? (const void *)scan_f32_kernel<K, N, M, 8>                        \
                         : (const void *)scan_f32_kernel<K, N, M, 4>;
    KNN_SCAN_INSTANCES(X)
#undef X
    return nullptr;
}

// Resident blocks per CU of one scan instance (VGPR and LDS limits).
int scan_blocks_per_cu(int ks, int bmode, int waves, int kp, int oct) {
    if (waves != 4 && waves != 8 && waves != 1) return 0;
    const void *fn = scan_fn(ks, bmode, waves);
    if (!fn) return 0;
    const size_t lds = scan_block_lds(kp, waves, bmode, oct);
    if (lds > 65536 &&
        hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return 0;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, waves * kWave, lds) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

hipError_t launch_scan(const ScanArgs &a, int ks, int bmode, int waves, hipStream_t s) {
    if (waves < 1 || waves > 8) return hipErrorInvalidValue;
    if (bmode == kBLds && a.oct % 4) return hipErrorInvalidValue;  // whole rings of 4 octets
    const void *fn = scan_fn(ks, bmode, waves);
    if (!fn) return hipErrorInvalidValue;
    const size_t lds = scan_block_lds(a.kp, waves, bmode, a.oct);
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const int blocks = a.tasks ? a.ntasks : a.nqb * a.nslices;
#define X(K, N, M)                                                                             \
    if (ks == K && bmode == M) {                                                               \
        if (waves > 4)                                                                         \
            hipLaunchKernelGGL((scan_f32_kernel<K, N, M, 8>), dim3(blocks), dim3(waves * kWave), \
                               lds, s, a);                                                     \
        else                                                                                   \
            hipLaunchKernelGGL((scan_f32_kernel<K, N, M, 4>), dim3(blocks), dim3(waves * kWave), \
                               lds, s, a);                                                     \
        return hipGetLastError();                                                              \
    }
    KNN_SCAN_INSTANCES(X)
#undef X
    return hipErrorInvalidValue;
}

// bfloat16 hi/lo layout of the precision scan: for tile t of 32 points and
// 16-dimension chunk c, [hi | lo] x 64 lanes x 8 bfloat16; lane m + 32h holds
// point 32t+m, dimensions 16c+8h .. +7.  hi = bf16(x), lo = bf16(x - hi),
// both rounded to nearest even.  Octet j of a point is chunk j/2, half j%2.
__device__ inline unsigned short bf16_bits(float x) {
    const unsigned u = __float_as_uint(x);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return (unsigned short)((u >> 16) | 0x40u);
    return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline void write_bf3(unsigned short *base, int tile, int nch, int j, int m,
                                 const float *v) {
    unsigned short hi[8], lo[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        hi[e] = bf16_bits(v[e]);
        lo[e] = bf16_bits(v[e] - __uint_as_float((unsigned)hi[e] << 16));
    }
    const int c = j >> 1, h = j & 1;
    unsigned short *dst = base + ((((long long)tile * nch + c) * 2) * kWave + m + 32 * h) * 8;
    uint4 ph, pl;
    ph.x = hi[0] | ((unsigned)hi[1] << 16);
    ph.y = hi[2] | ((unsigned)hi[3] << 16);
    ph.z = hi[4] | ((unsigned)hi[5] << 16);
    ph.w = hi[6] | ((unsigned)hi[7] << 16);
    pl.x = lo[0] | ((unsigned)lo[1] << 16);
    pl.y = lo[2] | ((unsigned)lo[3] << 16);
    pl.z = lo[4] | ((unsigned)lo[5] << 16);
    pl.w = lo[6] | ((unsigned)lo[7] << 16);
    *reinterpret_cast<uint4 *>(dst) = ph;
    *reinterpret_cast<uint4 *>(dst + kWave * 8) = pl;
}

// bfloat16 hi/lo tiles of n_pad points given as point-major rows (row stride
// ld, a multiple of 4, pads zero; dimensions past ld read as zero): one
// thread per (tile, chunk, lane).
__global__ __launch_bounds__(256) void pack_bf3_rows_kernel(const float *pm, int ld, int n_pad,
                                                            int nch, unsigned short *out) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long long)(n_pad / kRefsPerTile) * nch * kWave) return;
    const int lane = (int)(g % kWave);
    const long long tc = g / kWave;
    const int c = (int)(tc % nch);
    const int tile = (int)(tc / nch);
    const int m = lane & 31, h = lane >> 5;
    const float *row = pm + ((long long)tile * kRefsPerTile + m) * ld;
    float v[8];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const int d4 = 16 * c + 8 * h + 4 * p;
        const float4 x = d4 < ld ? *reinterpret_cast<const float4 *>(row + d4)
                                 : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        v[4 * p + 0] = x.x;
        v[4 * p + 1] = x.y;
        v[4 * p + 2] = x.z;
        v[4 * p + 3] = x.w;
    }
    write_bf3(out, tile, nch, 2 * c + h, m, v);
}

hipError_t launch_pack_bf3_rows(const float *pm, int ld, int n_pad, int nch, unsigned short *out,
                                hipStream_t s) {
    if (n_pad % kRefsPerTile || ld % 4 || nch < 1) return hipErrorInvalidValue;
    const long long threads = (long long)(n_pad / kRefsPerTile) * nch * kWave;
    if (threads == 0) return hipSuccess;
    hipLaunchKernelGGL(pack_bf3_rows_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
                       s, pm, ld, n_pad, nch, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Norms pass: squared norms, the maximum reference norm, the point-major copy
// used by the exact re-rank, and (references only) the packed tile layout
// read by the scan.  Points in [n, n_pad) get zero tiles.  Each point's norm
// is an FP32 fused multiply-add sum in dimension order.  Replaces
// code/knncuda.cu:217-227.
// ---------------------------------------------------------------------------
constexpr int kNormPts = 64;  // points per workgroup of the norms pass
__device__ __forceinline__ void norms_pack_points(const NormArgs &a, int i0) {
    // 64 points per workgroup; 32 dimensions (4 octets) at a time go through
    // LDS so that the column reads, the packed-tile writes and the
    // point-major row writes are all contiguous.
    constexpr int kPts = kNormPts, kDims = 32;
    __shared__ float tile[kDims][kPts + 1];
    const int t = threadIdx.x;
    float s = 0.0f;  // thread t < 64: squared norm of point i0 + t, summed in dimension order
    const int ngrp = (a.oct + 3) / 4;
    // rows starting on 16 bytes (pitch a multiple of 4, aligned base) are read
    // four points at a time; i0 is a multiple of 64
    const bool vec = (a.ld & 3) == 0 && (reinterpret_cast<unsigned long long>(a.p) & 15) == 0;
    for (int g = 0; g < ngrp; ++g) {
        if (vec) {
#pragma unroll
            for (int e = 0; e < (kDims * kPts) / (4 * 256); ++e) {
                const int idx = t + 256 * e;
                const int dl = idx / (kPts / 4), p = 4 * (idx % (kPts / 4));
                const int d = kDims * g + dl, i = i0 + p;
                float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (d < a.dim) {
                    const float *row = a.p + (long long)d * a.ld;
                    if (i + 3 < a.n) {
                        v = *reinterpret_cast<const float4 *>(row + i);
                    } else {
                        v.x = i + 0 < a.n ? row[i + 0] : 0.0f;
                        v.y = i + 1 < a.n ? row[i + 1] : 0.0f;
                        v.z = i + 2 < a.n ? row[i + 2] : 0.0f;
                        v.w = i + 3 < a.n ? row[i + 3] : 0.0f;
                    }
                }
                tile[dl][p + 0] = v.x;
                tile[dl][p + 1] = v.y;
                tile[dl][p + 2] = v.z;
                tile[dl][p + 3] = v.w;
            }
        } else {
#pragma unroll
            for (int e = 0; e < (kDims * kPts) / 256; ++e) {
                const int idx = t + 256 * e;
                const int dl = idx / kPts, p = idx % kPts;
                const int d = kDims * g + dl, i = i0 + p;
                tile[dl][p] = (i < a.n && d < a.dim) ? a.p[(long long)d * a.ld + i] : 0.0f;
            }
        }
        __syncthreads();
        if (t < kPts) {
#pragma unroll
            for (int dl = 0; dl < kDims; ++dl) {
                const float v = tile[dl][t];
                if (8 * (4 * g) + dl < 8 * a.oct) s = fmaf(v, v, s);
            }
        }
        // packed FP32 tiles: (octet j, tile, lane m + 32h) <- dims 8j + 4h .. +3 of point m
        if (a.packed) {
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int idx = t + 256 * r;
                const int jj = idx / 128, tl = (idx / 64) % 2, lane = idx % 64;
                const int m = lane % 32, h = lane / 32;
                const int j = 4 * g + jj, i = i0 + 32 * tl;
                if (j < a.oct && i < a.n_pad) {
                    const float4 v = make_float4(tile[8 * jj + 4 * h + 0][32 * tl + m],
                                                 tile[8 * jj + 4 * h + 1][32 * tl + m],
                                                 tile[8 * jj + 4 * h + 2][32 * tl + m],
                                                 tile[8 * jj + 4 * h + 3][32 * tl + m]);
                    *reinterpret_cast<float4 *>(
                        a.packed + (((long long)(i / kRefsPerTile) * a.oct + j) * kWave + lane) * 4) = v;
                }
            }
        }
        // bfloat16 hi/lo tiles: octet j of point m
        if (a.packed_bf) {
            const int jj = t / 64, tl = (t / 32) % 2, m = t % 32;
            const int j = 4 * g + jj, i = i0 + 32 * tl;
            if (j < a.oct && i < a.n_pad) {
                float v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = tile[8 * jj + u][32 * tl + m];
                write_bf3(a.packed_bf, i / kRefsPerTile, a.nch_bf, j, m, v);
            }
        }
        // point-major rows: 8 threads write a point's 32 dimensions (128 bytes)
        if (a.pm) {
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int idx = t + 256 * r;
                const int p = idx / 8, q4 = idx % 8;
                const int d = kDims * g + 4 * q4, i = i0 + p;
                if (i < a.n && d < a.pm_ld && 4 * q4 < 8 * (a.oct - 4 * g))
                    *reinterpret_cast<float4 *>(a.pm + (long long)i * a.pm_ld + d) =
                        make_float4(tile[4 * q4 + 0][p], tile[4 * q4 + 1][p], tile[4 * q4 + 2][p],
                                    tile[4 * q4 + 3][p]);
            }
        }
        __syncthreads();
    }
    // octets past the dimension, up to the precision scan's chunk count, are zero
    if (a.packed_bf && t < kPts) {
        const int i = i0 + t;
        if (i < a.n_pad) {
            const float z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int j = a.oct; j < 2 * a.nch_bf; ++j)
                write_bf3(a.packed_bf, i / kRefsPerTile, a.nch_bf, j, i % kRefsPerTile, z);
        }
    }
    if (t < kPts) {
        const int i = i0 + t;
        if (i < a.n) a.out[i] = s;
        // NaN or infinite coordinates (or squares that overflow) leave a norm
        // that is not a finite float: the search reports the input instead
        // of ordering NaN keys
        if (a.nonfinite && i < a.n && !(s <= 3.402823466e38f)) atomicOr(a.nonfinite, 1u);
        if (a.maxbits) {
            unsigned mbits = i < a.n ? __float_as_uint(s) : 0u;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1)
                mbits = max(mbits, (unsigned)__shfl_xor((int)mbits, off));
            if (t == 0) atomicMax(a.maxbits, mbits);
        }
    }
}

__global__ __launch_bounds__(256) void norms_pack_kernel(NormArgs a) {
    norms_pack_points(a, blockIdx.x * kNormPts);
}

// Many point sets in one launch: workgroup w takes points wg[w].y .. +63 of
// set wg[w].x, written at that set's start in the outputs of `a`.
__global__ __launch_bounds__(256) void norms_pack_jobs_kernel(NormArgs a, const NormJob *__restrict__ jobs,
                                                              const int2 *__restrict__ wg) {
    const int2 w = wg[blockIdx.x];
    const NormJob j = jobs[w.x];
    NormArgs b = a;
    b.p = j.p;
    b.ld = j.ld;
    b.n = j.n;
    b.n_pad = j.n_pad;
    b.out = a.out + j.start;
    b.pm = a.pm ? a.pm + (long long)j.start * a.pm_ld : nullptr;
    b.packed = a.packed ? a.packed + (long long)j.start * 8 * a.oct : nullptr;
    b.packed_bf = a.packed_bf ? a.packed_bf + (long long)(j.start / kRefsPerTile) * a.nch_bf * 2 * kWave * 8
                              : nullptr;
    norms_pack_points(b, w.y);
}

hipError_t launch_norms(const NormArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(norms_pack_kernel, dim3((a.n_pad + kNormPts - 1) / kNormPts), dim3(256), 0, s, a);
    return hipGetLastError();
}

// jobs / wg: device arrays (wg: nwg (set, first point) pairs, the first point
// a multiple of 64 below the set's n_pad).
hipError_t launch_norms_jobs(const NormArgs &a, const NormJob *jobs, const int2 *wg, int nwg,
                             hipStream_t s) {
    if (nwg <= 0) return hipSuccess;
    hipLaunchKernelGGL(norms_pack_jobs_kernel, dim3(nwg), dim3(256), 0, s, a, jobs, wg);
    return hipGetLastError();
}

}  // namespace knnhip
