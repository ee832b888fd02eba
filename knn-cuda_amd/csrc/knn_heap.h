// Device helpers shared by the scan kernels: accumulator row map, LDS
// max-heaps of (key, index) pairs and the per-query LDS locks that guard them.
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

}  // namespace knnhip
