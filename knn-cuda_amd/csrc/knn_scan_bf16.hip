// Precision-form distance scan: the FP32 scan's key ||r||^2 - 2 r.q computed
// with bfloat16 MFMA products (v_mfma_f32_32x32x16_bf16, FP32 accumulation)
// on a three-term split, fused with the same streaming top-kp selection.
//
// Split: every operand x (reference coordinate, or -2q) is x_hi + x_lo with
// x_hi = bf16(x) and x_lo = bf16(x - x_hi); the product r.q' is taken as
// r_hi.q_hi + r_hi.q_lo + r_lo.q_hi (three MFMAs per 16 dimensions).  With
// u = 2^-8 each dropped or rounded part is below 3.01 u^2 |r_i q'_i|, so the
// key stays within err_q ||q||^2 + err_r max||r||^2 of the exact one (the
// coefficients are set by the host and checked by the re-rank certificate,
// which re-ranks every candidate with the reference's FP32 formula).  The
// result is therefore the same as the FP32 path's, bit for bit.
//
// Mapping: a block holds 128 queries (four sets of 32) and 8 waves; waves w
// and w+4 share query set w%4 (its heaps, under the per-query LDS locks)
// and take tiles {0,1} / {2,3} of every 128-reference step.  Each 16-dim
// chunk of a step (4 tiles x hi/lo x 64 lanes x 16 bytes = 8 KiB) is loaded
// once per block, one 16-byte load per thread, into a four-slot LDS ring,
// so reference traffic is a quarter of the FP32 scan's (32 queries per
// block).  Global loads run PF chunks ahead in registers; each wave reads a
// chunk's operands from LDS one chunk ahead of its MFMAs; one barrier per
// two chunks hands slots between the loading threads and the computing
// waves.  Chunk g + 3 is stored during chunk g into slot (g + 3) % 4, and
// chunk g + 1 is read during chunk g.  With a barrier after every odd chunk
// and one after the prologue's read of chunk 0:
//   read-after-write: chunk g + 1 was stored in iteration g - 2 (or in the
//     prologue), and a barrier lies between (after g - 1 or g - 2, whichever
//     is odd);
//   write-after-read: slot (g + 3) % 4 last held chunk g - 1, read in
//     iteration g - 2 (chunk 0: in the prologue), and a barrier lies between
//     that read and iteration g's store (for g = 1 it is the prologue's).
#include "knn_common.h"
#include "knn_heap.h"

namespace knnhip {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef unsigned u32x4_t __attribute__((ext_vector_type(4)));

constexpr int kBfWaves = 8;
constexpr int kBfQueries = 128;
constexpr int kBfQueueCap = 384;  // >= 4 keys x 64 lanes between drain checks, plus slack
constexpr size_t kBfQueueBytes = (size_t)kBfQueueCap * 3 * sizeof(float);
constexpr int kBfChunkBytes = 4 * 2 * kWave * 16;  // one 16-dim chunk of a 128-reference step
constexpr int kBfRing = 4;                         // LDS slots

__host__ __device__ constexpr size_t bf_heap_bytes(int kp) {
    return (((size_t)kBfQueries * heap_stride(kp) * 2 + kBfQueries) * sizeof(float) + 15) &
           ~(size_t)15;
}

size_t scan_bf3_lds_bytes(int kp) {
    return bf_heap_bytes(kp) + (size_t)kBfWaves * kBfQueueBytes + kBfRing * (size_t)kBfChunkBytes;
}

// Workgroup barrier for LDS hand-offs: waits for this wave's LDS operations
// only, so the global loads already in flight (the prefetch) keep running;
// __syncthreads() would also wait for every outstanding global load.  The
// memory clobber keeps the compiler from moving LDS accesses across it.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

__device__ inline bf16x8_t split_hi_lo(const float *x, bf16x8_t &lo) {
    bf16x8_t hi;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const __bf16 h = (__bf16)x[e];
        hi[e] = h;
        lo[e] = (__bf16)(x[e] - (float)h);
    }
    return hi;
}

// NCH: 16-dimension chunks (padded dim = 16 NCH); PF: chunks loaded ahead
// (divides NCH, so the register ring index is static).  BF3: bfloat16
// three-term products on the hi/lo tiles; else FP32 products
// (v_mfma_f32_32x32x2_f32) on the FP32 packed tiles, whose two octets per
// 16-dim chunk sit where the hi and lo halves do.
template <int NCH, int PF, bool BF3>
__global__ __launch_bounds__(kBfWaves * kWave, 1) void scan_lds_kernel(ScanArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    static_assert(NCH % PF == 0, "ring position must repeat every step");
    static_assert(NCH % 2 == 0, "one barrier per two chunks, ending every step");
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 31, h = lane >> 5;
    const int qset = wave & 3, half = wave >> 2;

    const int t = blockIdx.x;
    const int slice = t / a.nqb, qb = t - slice * a.nqb;
    const int qfirst = qb * kBfQueries;
    const int nq = min(kBfQueries, a.Q - qfirst);
    if (nq <= 0) return;  // uniform over the block
    const int r_begin = slice * a.slice_len;
    const int r_end = min(a.R, r_begin + a.slice_len);
    const int qlocal = qset * kQueriesPerWave + col;
    const bool qvalid = qlocal < nq;
    const int q = qfirst + (qvalid ? qlocal : 0);

    // LDS: heaps + locks (128 queries), 8 candidate queues, the chunk ring
    const int kp = a.kp, kps = heap_stride(kp);
    float *hv = reinterpret_cast<float *>(smem);
    int *hi = reinterpret_cast<int *>(smem + (size_t)kBfQueries * kps * sizeof(float));
    int *lock = hi + (size_t)kBfQueries * kps;
    unsigned char *qbase = smem + bf_heap_bytes(kp) + (size_t)wave * kBfQueueBytes;
    float *qv = reinterpret_cast<float *>(qbase);
    int *qr = reinterpret_cast<int *>(qbase + kBfQueueCap * sizeof(float));
    int *qc = reinterpret_cast<int *>(qbase + kBfQueueCap * 2 * sizeof(float));
    unsigned char *ring = smem + bf_heap_bytes(kp) + (size_t)kBfWaves * kBfQueueBytes;
    for (int e = threadIdx.x; e < kBfQueries * kps; e += blockDim.x) {
        hv[e] = INFINITY;
        hi[e] = -1;
    }
    if (threadIdx.x < kBfQueries) lock[threadIdx.x] = 0;

    // B operands, per chunk: bf16 hi/lo of -2q (lane: query col, dims
    // 16c + 8h ..), or FP32 -2q (dims 16c + 8o + 4h + u, octet o, k-step u)
    bf16x8_t bh[BF3 ? NCH : 1], bl[BF3 ? NCH : 1];
    float bq[BF3 ? 1 : NCH][8];
    {
        const float *qcol = a.query + q;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            float x[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int d = BF3 ? 16 * c + 8 * h + e : 16 * c + 8 * (e >> 2) + 4 * h + (e & 3);
                const float v = qcol[(long long)min(d, a.dim - 1) * a.ldq];
                x[e] = d < a.dim ? -2.0f * v : 0.0f;
            }
            if constexpr (BF3) {
                bh[c] = split_hi_lo(x, bl[c]);
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) bq[c][e] = x[e];
            }
        }
    }

    float *myv = hv + qlocal * kps;
    float theta = INFINITY;
    float published = INFINITY;
    unsigned *theta_q = a.theta + q;
    int qcount = 0;  // wave-uniform
    const unsigned long long lanes_below = (1ull << lane) - 1ull;
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
            // wave-uniform retry loop (see the FP32 scan): one attempt per lane
            // per iteration, a lock released in the iteration that took it
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

    // Reference chunks: thread t loads 16 bytes of (tile t/128, hi/lo, lane)
    // of each chunk; the buffer resource covers the slice's tiles only, so
    // loads past the slice return zeros.
    constexpr long long tile_bytes = (long long)NCH * 2 * kWave * 16;
    const int nsteps = (r_end - r_begin + kRefsPerStep - 1) / kRefsPerStep;
    const long long valid_tiles =
        max(0, min(nsteps * kRefsPerStep, a.r_pad - r_begin)) / kRefsPerTile;
    const unsigned char *pbase = BF3 ? reinterpret_cast<const unsigned char *>(a.packed_bf)
                                     : reinterpret_cast<const unsigned char *>(a.packed);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<unsigned char *>(pbase) + (long long)(r_begin / kRefsPerTile) * tile_bytes,
        (short)0, (int)(unsigned)(valid_tiles * tile_bytes), 0x00020000);
    const int ltile = threadIdx.x >> 7, lhl = (threadIdx.x >> 6) & 1, lln = threadIdx.x & 63;
    const unsigned my_off = (unsigned)(ltile * tile_bytes + lhl * kWave * 16 + lln * 16);
    // chunk st * NCH + j (j >= 0 may pass NCH) lives at (st + j / NCH) * 4
    // tiles + (j % NCH) * 2 KiB.  32-bit offsets: the host keeps a slice's
    // tiles within 2^31 bytes and loads run at most two steps past the slice
    // (they read zeros there: the offset is past the resource's size).
    constexpr unsigned step_bytes = 4u * (unsigned)tile_bytes;
    auto chunk_off = [&](int st, int j) -> unsigned {
        return my_off + (unsigned)(st + j / NCH) * step_bytes + (unsigned)(j % NCH) * (2u * kWave * 16);
    };
    // Ring invariant at the top of chunk g: chunks g .. g+2 sit in slots
    // x % 4 (chunks g+1 and g+2 complete since the last barrier when g is
    // even, chunk g+2 possibly still being stored when g is odd),
    // pre[x % PF] holds chunk x for g+3 <= x < g+3+PF, and this wave's
    // operands of chunk g are in cur[].  Prologue: chunks 0..2 into slots
    // 0..2.
#pragma unroll
    for (int j = 0; j < 3; ++j)
        *reinterpret_cast<u32x4_t *>(ring + j * kBfChunkBytes + threadIdx.x * 16) =
            __builtin_amdgcn_raw_buffer_load_b128(rs, (int)chunk_off(0, j), 0, 0);
    u32x4_t pre[PF];
#pragma unroll
    for (int j = 0; j < PF; ++j)
        pre[(3 + j) % PF] = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)chunk_off(0, 3 + j), 0, 0);
    __syncthreads();
    const int my_t0 = 2 * half;  // my tiles: my_t0, my_t0 + 1
    auto read_ops = [&](int slotidx, u32x4_t *ops) {
        const unsigned char *slot = ring + slotidx * kBfChunkBytes;
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int hl = 0; hl < 2; ++hl)
                ops[2 * mt + hl] =
                    *reinterpret_cast<const u32x4_t *>(slot + (((my_t0 + mt) * 2 + hl) * kWave + lane) * 16);
    };
    u32x4_t cur[4];
    read_ops(0, cur);
    // Slot 0 is rewritten (with chunk 4) in iteration 1, before the loop's
    // first barrier: every wave's read of chunk 0 above must be done first.
    // Without this barrier a wave that reached iteration 1 early could
    // overwrite chunk 0 while a slower wave had yet to read it, and that
    // wave would then multiply chunk 4's tiles as chunk 0's: wrong keys for
    // its first step's references, which the certificate cannot see.
    lds_barrier();

    for (int st = 0; st < nsteps; ++st) {
        const int r0 = r_begin + st * kRefsPerStep;
        f32x16_t acc[2];
        // start from ||r||^2 (rows 8(i/4) + 4h + i%4 of each tile)
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int g4 = 0; g4 < 4; ++g4) {
                const float4 nr = *reinterpret_cast<const float4 *>(
                    a.rnorm + r0 + kRefsPerTile * (my_t0 + mt) + 8 * g4 + 4 * h);
                acc[mt][4 * g4 + 0] = nr.x;
                acc[mt][4 * g4 + 1] = nr.y;
                acc[mt][4 * g4 + 2] = nr.z;
                acc[mt][4 * g4 + 3] = nr.w;
            }
        const unsigned tq = __hip_atomic_load(theta_q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const int g = st * NCH + c;
            // next chunk's operands (slot (g+1)%4, stored in iteration g-2,
            // complete since the last barrier)
            u32x4_t nxt[4];
            read_ops((g + 1) % kBfRing, nxt);
            // chunk g+3 into slot (g+3)%4 (it held chunk g-1, which every wave
            // read in iteration g-2, before the last barrier); refill its
            // register
            {
                const int rslot = (c + 3) % PF;
                *reinterpret_cast<u32x4_t *>(ring + ((g + 3) % kBfRing) * kBfChunkBytes + threadIdx.x * 16) =
                    pre[rslot];
                pre[rslot] = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)chunk_off(st, c + 3 + PF), 0, 0);
            }
            if constexpr (BF3) {
                const bf16x8_t a0h = __builtin_bit_cast(bf16x8_t, cur[0]);
                const bf16x8_t a0l = __builtin_bit_cast(bf16x8_t, cur[1]);
                const bf16x8_t a1h = __builtin_bit_cast(bf16x8_t, cur[2]);
                const bf16x8_t a1l = __builtin_bit_cast(bf16x8_t, cur[3]);
                acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0h, bh[c], acc[0], 0, 0, 0);
                acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1h, bh[c], acc[1], 0, 0, 0);
                acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0h, bl[c], acc[0], 0, 0, 0);
                acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1h, bl[c], acc[1], 0, 0, 0);
                acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0l, bh[c], acc[0], 0, 0, 0);
                acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1l, bh[c], acc[1], 0, 0, 0);
            } else {
                // octet o of tile mt is cur[2 mt + o]; k-step u of an octet reads
                // dimension 8o + 4h + u on both operands
#pragma unroll
                for (int o = 0; o < 2; ++o)
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        acc[0] = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(cur[o][u]),
                                                                      bq[c][4 * o + u], acc[0], 0, 0, 0);
                        acc[1] = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(cur[2 + o][u]),
                                                                      bq[c][4 * o + u], acc[1], 0, 0, 0);
                    }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) cur[j] = nxt[j];
            // every second chunk (NCH is even): chunks g+2 and g+3 stored,
            // everyone past chunk g's and g+1's reads
            if (c & 1) lds_barrier();
        }

        // Epilogue (as the FP32 scan): keys below the query's threshold go
        // to the wave's queue, drained into the shared heaps.
        theta = fminf(fminf(theta, ordered_to_float(tq)), shared_root(myv));
        bool cand = false;
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int i = 0; i < 16; ++i) cand |= acc[mt][i] < theta;
        cand = cand && qvalid;
        if (__ballot(cand)) {
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) {
#pragma unroll
                for (int g = 0; g < 16 / kQueueGroup; ++g) {
                    // theta only falls, so a skipped group holds no key below it
                    if (__ballot(qvalid && group_min(acc[mt], g) < theta) == 0) continue;
                    if (qcount > kBfQueueCap - kQueueGroup * kWave) {
                        drain();
                        theta = fminf(theta, shared_root(myv));
                    }
#pragma unroll
                    for (int e = 0; e < kQueueGroup; ++e) {
                        const int i = kQueueGroup * g + e;
                        const int r = r0 + kRefsPerTile * (my_t0 + mt) + acc_row(i, h);
                        const bool take = qvalid && acc[mt][i] < theta && r < r_end;
                        const unsigned long long m = __ballot(take);
                        if (m == 0) continue;
                        if (take) {
                            const int pos = qcount + __popcll(m & lanes_below);
                            qv[pos] = acc[mt][i];
                            qr[pos] = r;
                            qc[pos] = qlocal;
                        }
                        qcount += __popcll(m);
                    }
                }
            }
            // drain early: the heap root that filters the next steps' keys
            // (and the published threshold) falls only at a drain (32 keys
            // measured best against 16, 64, 128 and half the queue)
            if (qcount >= kWave / 2) {
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

    // heaps of the block's queries: rows slice * (nqb * 128) + query
    __syncthreads();
    const long long obase = ((long long)slice * a.nqb * kBfQueries + qfirst) * kp;
    for (int e = threadIdx.x; e < nq * kp; e += blockDim.x) {
        const int row = e / kp, j = e - row * kp;
        a.part_val[obase + e] = hv[row * kps + j];
        a.part_idx[obase + e] = hi[row * kps + j];
    }
}

#define KNN_LDS_INSTANCES(X) \
    X(2, 2)                  \
    X(4, 4)                  \
    X(6, 3)                  \
    X(8, 4)                  \
    X(12, 4)

int scan_bf3_nch(int dim) {
    const int n = (dim + 15) / 16;
    for (int c : {2, 4, 6, 8, 12})
        if (c >= n) return c;
    return 0;  // wider dimensions use the 32-query scan (the B operands would not fit in VGPRs)
}

// bf3: bfloat16 three-term products (packed_bf), else FP32 (packed).
hipError_t launch_scan_lds(const ScanArgs &a, int nch, bool bf3, hipStream_t s) {
    const size_t lds = scan_bf3_lds_bytes(a.kp);
    const int blocks = a.nqb * a.nslices;
#define X(N, P)                                                                                 \
    if (nch == N) {                                                                             \
        const void *fn = bf3 ? (const void *)scan_lds_kernel<N, P, true>                         \
                             : (const void *)scan_lds_kernel<N, P, false>;                       \
        if (lds > 65536) {                                                                      \
            hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,   \
                                               (int)lds);                                       \
            if (e != hipSuccess) return e;                                                      \
        }                                                                                       \
        if (bf3)                                                                                \
            hipLaunchKernelGGL((scan_lds_kernel<N, P, true>), dim3(blocks), dim3(kBfWaves * kWave), \
                               lds, s, a);                                                      \
        else                                                                                    \
            hipLaunchKernelGGL((scan_lds_kernel<N, P, false>), dim3(blocks), dim3(kBfWaves * kWave), \
                               lds, s, a);                                                      \
        return hipGetLastError();                                                               \
    }
    KNN_LDS_INSTANCES(X)
#undef X
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Task form of the precision scan, for task lists whose tasks hold up to 32
// queries (block-index and batched searches): the FP32 scan's block shape
// (four waves on one task's 32 queries, wave w taking reference steps w,
// w+4, ..., all feeding one LDS heap per query) with the three-term bfloat16
// products.  The hi + lo tiles hold as many bytes per reference as the FP32
// tiles, and a 32 x 32 tile costs 3 x 32 MFMA cycles per 16 dimensions
// instead of 8 x 64.  Keys start from ||r||^2 and add the same products in
// the same chunk order as the 128-query form, so the same error bound holds.
// NB: chunks in the register ring (loads run NB-1 chunks ahead).
// IND: every wave is a task of its own (tasks waves * block + wave), with its
// own heaps and queue in its own LDS region and no workgroup barrier: the
// tasks next to each other in the list, which hold the same block's next
// query groups, run on one CU and read the block's tiles once from HBM.
// MERGE (one-wave tasks of up to kMergeQueries queries, kp <= kMergeMaxKp):
// sorted per-query lists in place of the heaps and the queue; a step's
// candidates join the lists in one merge (ball tasks, whose few queries each
// see tens of candidates a step).  The task's rows start at its list offset.
// ---------------------------------------------------------------------------
template <int NCH, int NB, bool IND, bool MERGE>
__global__ __launch_bounds__(256, 2) void scan_bf3_task_kernel(ScanArgs a) {
    extern __shared__ __align__(16) unsigned char smem_all[];
    static_assert(NCH % NB == 0 && NB >= 2, "ring position must repeat every step");
    static_assert(IND || !MERGE, "merged lists are per wave");
    const int nwaves = blockDim.x >> 6;
    const int wave_id = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 31, h = lane >> 5;
    // within the task: IND tasks are one wave
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
    // columns past the task's queries compute query 0's keys and take nothing
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

    // B operands: bfloat16 hi / lo of -2q, lane (col, h) holding dimensions
    // 16c + 8h .. +7 of chunk c
    // (every chunk's pieces are loaded before any is used: one round trip)
    bf16x8_t bh[NCH], bl[NCH];
    float x[NCH][8];
    if (a.query_pm) {
        const float *row = a.query_pm + (long long)q * a.query_pm_ld;
#pragma unroll
        for (int c = 0; c < NCH; ++c)
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const float4 v = row_piece(row, 16 * c + 8 * h + 4 * p, a.query_pm_ld);
                x[c][4 * p + 0] = -2.0f * v.x;
                x[c][4 * p + 1] = -2.0f * v.y;
                x[c][4 * p + 2] = -2.0f * v.z;
                x[c][4 * p + 3] = -2.0f * v.w;
            }
    } else {
        const float *qcol = a.query + q;
#pragma unroll
        for (int c = 0; c < NCH; ++c)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int d = 16 * c + 8 * h + e;
                const float v = qcol[(long long)min(d, a.dim - 1) * a.ldq];
                x[c][e] = d < a.dim ? -2.0f * v : 0.0f;
            }
    }
#pragma unroll
    for (int c = 0; c < NCH; ++c) bh[c] = split_hi_lo(x[c], bl[c]);

    // A operands through one buffer resource over the task's tiles: loads
    // past them (the prefetch beyond a wave's last step) return zeros.
    constexpr long long kTileBytes = (long long)NCH * 2 * kWave * 16;
    const int span_refs = (r_end - r_begin + kRefsPerStep - 1) / kRefsPerStep * kRefsPerStep;
    const long long valid_tiles = max(0, min(span_refs, a.r_pad - r_begin)) / kRefsPerTile;
    const __amdgpu_buffer_rsrc_t rs =
        uniform_buffer_rsrc(a.packed_bf + (long long)(r_begin / kRefsPerTile) * (kTileBytes / 2),
                            (unsigned)(valid_tiles * kTileBytes));
    auto lane_offsets = [&](int r_start, unsigned *vo) {
        const long long t0 = (r_start - r_begin) / kRefsPerTile;
#pragma unroll
        for (int mt = 0; mt < kMTiles; ++mt)
            vo[mt] = (unsigned)min((t0 + mt) * kTileBytes + lane * 16, 0xFFFFFFF0ll);
    };
    u32x4_t ring[NB][kMTiles][2];
    auto issue = [&](u32x4_t(&buf)[kMTiles][2], int c, const unsigned *vo) {
#pragma unroll
        for (int mt = 0; mt < kMTiles; ++mt)
#pragma unroll
            for (int hl = 0; hl < 2; ++hl)
                buf[mt][hl] = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)vo[mt],
                                                                   (2 * c + hl) * kWave * 16, 0);
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
        // start from ||r||^2: element i of tile mt is reference
        // r0 + 32 mt + 8 (i / 4) + 4h + i % 4
        if (r0 + kRefsPerStep <= a.R) {
#pragma unroll
            for (int mt = 0; mt < kMTiles; ++mt)
#pragma unroll
                for (int g4 = 0; g4 < 4; ++g4) {
                    const float4 nr = *reinterpret_cast<const float4 *>(
                        a.rnorm + r0 + kRefsPerTile * mt + 8 * g4 + 4 * h);
                    acc[mt][4 * g4 + 0] = nr.x;
                    acc[mt][4 * g4 + 1] = nr.y;
                    acc[mt][4 * g4 + 2] = nr.z;
                    acc[mt][4 * g4 + 3] = nr.w;
                }
        } else {
#pragma unroll
            for (int mt = 0; mt < kMTiles; ++mt)
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int r = r0 + kRefsPerTile * mt + acc_row(i, h);
                    acc[mt][i] = r < a.R ? a.rnorm[r] : INFINITY;
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
            // three products per tile; consecutive MFMAs target different
            // accumulators
#pragma unroll
            for (int mt = 0; mt < kMTiles; ++mt)
                acc[mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    __builtin_bit_cast(bf16x8_t, ring[c % NB][mt][0]), bh[c], acc[mt], 0, 0, 0);
#pragma unroll
            for (int mt = 0; mt < kMTiles; ++mt)
                acc[mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    __builtin_bit_cast(bf16x8_t, ring[c % NB][mt][0]), bl[c], acc[mt], 0, 0, 0);
#pragma unroll
            for (int mt = 0; mt < kMTiles; ++mt)
                acc[mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    __builtin_bit_cast(bf16x8_t, ring[c % NB][mt][1]), bh[c], acc[mt], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int mt = 0; mt < kMTiles; ++mt) voff[mt] = voff_next[mt];

        // Epilogue (as the FP32 scan): keys below the freshest threshold go
        // to the wave's queue, drained into the shared heaps when it fills
        // (merged lists: to the lane's candidate slots, merged every step).
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
        // the step's smallest key against theta first; candidates are counted
        // only on steps that have one
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
                for (int i = 0; i < 16; ++i) {
                    // block-index tasks see dense candidates: every row is
                    // tested (a group test costs more than it skips there)
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
        __hip_atomic_fetch_min(theta_q, float_to_ordered(theta), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);

    // the task's heaps leave as rows task_row .. task_row + nq - 1
    task_sync();
    const long long obase = (long long)task_row * kp;
    for (int e = tid; e < nq * kp; e += tstride) {
        const int row = e / kp, j = e - row * kp;
        a.part_val[obase + e] = hv[row * kps + j];
        a.part_idx[obase + e] = hi[row * kps + j];
    }
}

#define KNN_BF3_TASK_INSTANCES(X) \
    X(2, 2)                       \
    X(4, 2)                       \
    X(6, 3)                       \
    X(8, 2)

// Chunks of the task form for a dimension, or 0: wider dimensions keep FP32
// products (twelve chunks of B operands would not fit in VGPRs beside the
// four accumulators and the load ring).
int scan_bf3_task_nch(int dim) {
    const int n = (dim + 15) / 16;
    for (int c : {2, 4, 6, 8})
        if (c >= n) return c;
    return 0;
}

// Task lists only (a.tasks, one block of `waves` waves per task), with the
// LDS of the FP32 scan's 32-query blocks of that many waves.  merge: one-wave
// tasks of up to kMergeQueries queries with merged lists (kp <= kMergeMaxKp).
hipError_t launch_scan_bf3_task(const ScanArgs &a, int nch, int waves, size_t lds, hipStream_t s,
                                int tasks_per_group, bool merge) {
    if (!a.tasks || !a.packed_bf) return hipErrorInvalidValue;
    if (tasks_per_group > 1 && (waves != 1 || tasks_per_group > 4)) return hipErrorInvalidValue;
    if (merge && (waves != 1 || tasks_per_group < 1 || a.kp > kMergeMaxKp)) return hipErrorInvalidValue;
    const size_t lds_ind = (size_t)tasks_per_group * (merge ? merge_lds_bytes(a.kp) : scan_heap_bytes(a.kp) + kQueueBytes);
    const dim3 grid_ind((a.ntasks + tasks_per_group - 1) / tasks_per_group), block_ind(tasks_per_group * kWave);
#define LAUNCH_IND(N, B, M)                                                                     \
    {                                                                                           \
        const void *fi = (const void *)scan_bf3_task_kernel<N, B, true, M>;                     \
        if (lds_ind > 65536) {                                                                  \
            hipError_t e = hipFuncSetAttribute(fi, hipFuncAttributeMaxDynamicSharedMemorySize,   \
                                               (int)lds_ind);                                   \
            if (e != hipSuccess) return e;                                                      \
        }                                                                                       \
        hipLaunchKernelGGL((scan_bf3_task_kernel<N, B, true, M>), grid_ind, block_ind, lds_ind, s, a); \
        return hipGetLastError();                                                               \
    }
#define X(N, B)                                                                                 \
    if (nch == N) {                                                                             \
        if (merge) LAUNCH_IND(N, B, true)                                                       \
        if (tasks_per_group > 1) LAUNCH_IND(N, B, false)                                        \
        const void *fn = (const void *)scan_bf3_task_kernel<N, B, false, false>;                \
        if (lds > 65536) {                                                                      \
            hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,   \
                                               (int)lds);                                       \
            if (e != hipSuccess) return e;                                                      \
        }                                                                                       \
        hipLaunchKernelGGL((scan_bf3_task_kernel<N, B, false, false>), dim3(a.ntasks), dim3(waves * kWave), \
                           lds, s, a);                                                          \
        return hipGetLastError();                                                               \
    }
    KNN_BF3_TASK_INSTANCES(X)
#undef X
#undef LAUNCH_IND
    return hipErrorInvalidValue;
}

}  // namespace knnhip
