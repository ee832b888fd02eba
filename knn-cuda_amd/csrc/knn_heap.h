// Device helpers shared by the scan kernels: accumulator row map, LDS
// max-heaps of (key, index) pairs, the per-query LDS locks that guard them,
// the LDS layout of a 32-query scan block, and the merged sorted lists of
// one-wave ball tasks.
#pragma once

#include "knn_common.h"

namespace knnhip {

// Row within a 32-row tile of accumulator element i on lane half h of the
// 32x32 MFMA result (4 consecutive rows per group, groups 8 rows apart).
__device__ __forceinline__ constexpr int acc_row(int i, int h) {
    return ((i >> 2) << 3) + (h << 2) + (i & 3);
}

// Max-heap of kp (key, index) pairs: replace the root by a smaller key and
// restore the heap.  Returns the new root.
__device__ __forceinline__ float heap_replace_top(float *hv, int *hi, int kp, float v, int id) {
    int i = 0;
    while (true) {
        int c = 2 * i + 1;
        if (c >= kp) break;
        float cv = hv[c];
        if (c + 1 < kp) {
            float c2 = hv[c + 1];
            if (c2 > cv) {
                cv = c2;
                ++c;
            }
        }
        if (!(cv > v)) break;
        hv[i] = cv;
        hi[i] = hi[c];
        i = c;
    }
    hv[i] = v;
    hi[i] = id;
    return hv[0];
}

__host__ __device__ constexpr int heap_stride(int kp) { return kp | 1; }

// Heap root read that other waves of the block may be updating (an atomic
// load, so the compiler re-reads it on every use).
__device__ __forceinline__ float shared_root(float *hv) {
    return __uint_as_float(__hip_atomic_load(reinterpret_cast<unsigned *>(hv), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP));
}

// Per-query LDS lock.  LDS instructions of one wave execute in order, so a
// release store issued after the heap writes publishes them; the fences are
// compiler-ordering points only (no hardware wait is needed at wavefront
// scope).
__device__ __forceinline__ bool try_lock(int *l) {
    int expected = 0;
    const bool ok = __hip_atomic_compare_exchange_strong(l, &expected, 1, __ATOMIC_RELAXED,
                                                         __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_WORKGROUP);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return ok;
}
__device__ __forceinline__ void unlock(int *l) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __hip_atomic_store(l, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// LDS of a 32-query scan block: the shared heaps and locks, then one
// candidate queue (key, reference, query column) per wave.  Blocks of up to
// four waves queue up to 768 candidates a wave; eight-wave blocks, whose LDS
// is short, 384.
constexpr int kQueueCap = 768;
constexpr size_t kQueueBytes = (size_t)kQueueCap * 3 * sizeof(float);
__host__ __device__ constexpr int scan_queue_cap(int waves) { return waves > 4 ? 384 : kQueueCap; }
// A step's keys are queued four accumulator rows at a time.  A group whose
// four keys are at or above the threshold on every lane is skipped with one
// test; before a group that has a candidate, a queue with less than room for
// 4 x 64 keys is drained.
constexpr int kQueueGroup = 4;
template <class V> __device__ __forceinline__ float group_min(const V &acc, int g) {
    return fminf(fminf(acc[kQueueGroup * g], acc[kQueueGroup * g + 1]),
                 fminf(acc[kQueueGroup * g + 2], acc[kQueueGroup * g + 3]));
}
__host__ __device__ constexpr size_t scan_queue_bytes(int waves) {
    return (size_t)scan_queue_cap(waves) * 3 * sizeof(float);
}
__host__ __device__ constexpr size_t scan_heap_bytes(int kp) {
    return (((size_t)kQueriesPerWave * heap_stride(kp) * 2 + kQueriesPerWave) * sizeof(float) + 15) &
           ~(size_t)15;
}

// Inserts a wave's queued candidates into the block's shared heaps (heap c
// at hv / hi + c * kps, guarded by lock[c]).  The wave takes 64 entries at a
// time, one per lane; entries of one heap are serialised by its lock.  The
// retry loop is wave-uniform (its condition is a ballot) and each lane makes
// one attempt per iteration, releasing a lock in the iteration that took it:
// a per-lane spin loop would let the compiler move a winner's critical
// section past the loop while a losing lane of the same wave spins on that
// lock.  Entries no longer below their heap's root are dropped unlocked.
__device__ __forceinline__ void drain_queue(const float *qv, const int *qr, const int *qc,
                                            int qcount, float *hv, int *hi, int *lock, int kps,
                                            int kp, int lane) {
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
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Sorted candidate lists of one-wave ball tasks.  A ball task holds a few
// queries (2-6 on clustered data) and a step can bring tens of candidates per
// query: a heap takes them one sift-down at a time, behind its lock, with
// most of the wave's lanes idle.  Here each query keeps its kp best entries
// as a sorted list, and a whole batch of candidates joins it at once: every
// list entry and every candidate finds its rank in the union by counting
// (one lane an element), and the kp smallest are written back in rank order.
// A task holds up to kMergeQueries queries; between merges lane (col, h) of
// a column col < kMergeQueries holds up to kMergeCand candidates of its
// column in slot column h * kMergeQueries + col.
// ---------------------------------------------------------------------------
struct alignas(8) KeyRef {
    float k;
    int r;
};

__host__ __device__ constexpr size_t merge_lds_bytes(int kp) {
    return ((size_t)kMergeQueries * kp + (size_t)kMergeCand * kMergeSlots) * sizeof(KeyRef);
}

// Merges every column's pending candidates into its list (list q at
// lists + q * kp, nfill real entries, the rest (+inf, -1)).  cnt, nfill and
// thr are per lane: lane (col, h) holds its own candidate count and its
// column's fill and kp-th key (+inf while the list is short).  Wave-uniform
// control flow; the whole wave takes one query at a time: lane e holds list
// entry e (e < n) and lane j candidate j (half 0's, then half 1's, each in
// append order; at most 2 kMergeCand of them).  The rank of an element is
// its position in the union ordered by (key, union position), union
// position meaning the list's entries first, then the candidates:
//   list entry e      e + #{candidates t : key_t < key_e}   (the list is
//                     sorted, equal keys in union order, from the last merge)
//   candidate j       #{entries e : key_e <= key_j} (binary search of the
//                     sorted list) + #{candidates t : key_t < key_j, or
//                     key_t == key_j and t < j}
// so the loop runs over the candidates only, reading each one's key across
// the wave with readlane.  The ranks are a permutation of the union; the kp
// smallest land in order and the list stays sorted.
__device__ __forceinline__ float lane_key(float v, int j) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ __forceinline__ void merge_lists(KeyRef *lists, const KeyRef *cand, int kp, int nq, int lane,
                                            int col, int &cnt, int &nfill, float &thr) {
    static_assert(2 * kMergeCand <= kWave && kMergeMaxKp <= kWave, "one list entry and one candidate a lane");
    wave_sync();
    for (int q = 0; q < nq; ++q) {
        const int c0 = __builtin_amdgcn_readlane(cnt, q);
        const int c1 = __builtin_amdgcn_readlane(cnt, q + kWave / 2);
        const int c = c0 + c1;
        if (c == 0) continue;
        const int n = __builtin_amdgcn_readlane(nfill, q);
        KeyRef *list = lists + (size_t)q * kp;
        KeyRef le, cj;
        le.k = INFINITY;
        le.r = -1;
        cj.k = INFINITY;
        cj.r = -1;
        if (lane < n) le = list[lane];
        if (lane < c0)
            cj = cand[lane * kMergeSlots + q];
        else if (lane < c)
            cj = cand[(lane - c0) * kMergeSlots + kMergeQueries + q];
        // entries at or below the candidate: binary search of the sorted list
        int below = 0;
#pragma unroll
        for (int step = kWave; step > 0; step >>= 1)
            if (below + step <= n && list[below + step - 1].k <= cj.k) below += step;
        int rl = lane, rc = below;
        for (int t = 0; t < c; ++t) {
            const float y = lane_key(cj.k, t);
            rl += y < le.k ? 1 : 0;
            rc += (y < cj.k || (y == cj.k && t < lane)) ? 1 : 0;
        }
        wave_sync();  // every read of the list precedes the writes
        if (lane < n && rl < kp) list[rl] = le;
        if (lane < c && rc < kp) list[rc] = cj;
        wave_sync();
        const int m = n + c;
        const int nn = m < kp ? m : kp;
        if (col == q) {
            nfill = nn;
            if (nn == kp) thr = list[kp - 1].k;
        }
    }
    cnt = 0;
    wave_sync();
}

// Step-local cut of one query column.  A block's first steps can put most of
// a step's keys below the query's threshold (an empty heap's root is +inf),
// and every queued key costs a pass through the heap locks.  Bisection on
// the value of the column's keys below theta (lane halves h = 0, 1 hold the
// column's rows 8(i/4) + 4h + i%4 of every tile) returns c with at least kp
// of them below c: queueing only keys below c still brings kp of the step's
// keys to the heap, so every key left out is at or above the heap's final
// root.  Rows past the task must already read +inf.  Lanes with need ==
// false keep theta (the loop runs wave-wide).
constexpr int kCutIters = 8;
template <class V, int NT>
__device__ __forceinline__ float step_cut(const V (&acc)[NT], float theta, int kp, bool need) {
    float lo = INFINITY, hi = -INFINITY;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const float v = acc[t][i];
            const bool c = v < theta;
            lo = c ? fminf(lo, v) : lo;
            hi = c ? fmaxf(hi, v) : hi;
        }
    lo = fminf(lo, __shfl_xor(lo, 32));
    hi = nextafterf(fmaxf(hi, __shfl_xor(hi, 32)), INFINITY);  // every candidate is below hi
#pragma unroll 1
    for (int it = 0; it < kCutIters; ++it) {
        const float mid = 0.5f * lo + 0.5f * hi;
        int c = 0;
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int i = 0; i < 16; ++i) c += acc[t][i] < mid ? 1 : 0;
        c += __shfl_xor(c, 32);
        if (c >= kp)
            hi = mid;
        else
            lo = mid;
    }
    return need ? hi : theta;
}

// One step's candidates of a merged-list task (acc: the step's kMTiles
// 32 x 32 key tiles, rows outside the task already +inf): keys below theta
// (columns with many, as with a young list: below the step-local cut) go to
// the lane's candidate slots, merged into the lists when a tile could
// overflow the slots and at the step's end; theta then takes the column's
// kp-th key.
template <class V>
__device__ __forceinline__ void merge_step(const V (&acc)[kMTiles], float &theta, int kp, bool qvalid, int r0,
                                           int h, int col, int lane, int nq, KeyRef *lists, KeyRef *cand,
                                           int &mcount, int &mfill, float &mthr) {
    int ncand = 0;
#pragma unroll
    for (int mt = 0; mt < kMTiles; ++mt)
#pragma unroll
        for (int i = 0; i < 16; ++i) ncand += acc[mt][i] < theta ? 1 : 0;
    if (!qvalid) ncand = 0;
    const int npair = ncand + __shfl_xor(ncand, 32);
    float cut = theta;
    if (__ballot(npair >= 2 * kp)) cut = step_cut(acc, theta, kp, npair >= 2 * kp);
    const int mslot = h * kMergeQueries + col;
#pragma unroll
    for (int mt = 0; mt < kMTiles; ++mt) {
        int ct = 0;
#pragma unroll
        for (int i = 0; i < 16; ++i) ct += acc[mt][i] < cut ? 1 : 0;
        if (!qvalid) ct = 0;
        if (__ballot(ct > 0) == 0) continue;
        if (__ballot(mcount + ct > kMergeCand)) {
            merge_lists(lists, cand, kp, nq, lane, col, mcount, mfill, mthr);
            cut = fminf(cut, mthr);
        }
        if (qvalid) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                if (acc[mt][i] < cut) {
                    KeyRef e;
                    e.k = acc[mt][i];
                    e.r = r0 + kRefsPerTile * mt + acc_row(i, h);
                    cand[mcount * kMergeSlots + mslot] = e;
                    ++mcount;
                }
            }
        }
    }
    merge_lists(lists, cand, kp, nq, lane, col, mcount, mfill, mthr);
    theta = fminf(theta, mthr);
}

}  // namespace knnhip
