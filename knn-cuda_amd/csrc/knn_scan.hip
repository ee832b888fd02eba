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
// with v_mfma_f32_32x32x2_f32: A = references (32 rows per tile, read
// straight from the dim-major input, which is already the A-operand layout:
// lanes 0-31 hold dimension 2k, lanes 32-63 dimension 2k+1 of 32 consecutive
// points), B = -2 q resident in VGPRs for the whole scan, and one extra
// k-step that adds ||r||^2 (A = norm, B = 1).  The per-query constant
// ||q||^2 does not change the order and is added back by the re-rank stage.
//
// Selection: the accumulator holds 16 references of one query per lane.  A
// key below the query's current threshold theta (the heap root) is a
// candidate; the common case is one compare per key and a wave ballot.
// Candidates replace the heap root and sift down.  The two lane halves hold
// the same 32 queries, so they take turns and hand theta across with a
// shuffle.  Invariant used by the certificate: every reference not left in a
// heap has key >= that heap's final root.
#include "knn_common.h"

namespace knnhip {

typedef float f32x16 __attribute__((ext_vector_type(16)));

// Row within a 32-row tile of accumulator element i on lane half h of the
// 32x32 MFMA result (4 consecutive rows per group, groups 8 rows apart).
__device__ __forceinline__ constexpr int acc_row(int i, int h) {
    return ((i >> 2) << 3) + (h << 2) + (i & 3);
}

// Max-heap of kp (key, index) pairs: replace the root by a smaller key and
// restore the heap.  Returns the new root (the query's threshold).
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

template <int KS, bool MULTI>
__global__ __launch_bounds__(256, 2) void scan_f32_kernel(ScanArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int CH = (KS > 64) ? 4 : 8;  // k-steps per prefetch group
    static_assert(KS % CH == 0, "k-steps must be a multiple of the prefetch group");
    constexpr int NG = KS / CH;

    const int waves = blockDim.x >> 6;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int col = lane & 31;
    const int h = lane >> 5;

    int slice, qb;
    const int b = blockIdx.x;
    if (a.xcd_map) {
        // Consecutive block ids land on different XCDs; give every block of
        // one slice the same XCD so the slice is read from HBM once per L2.
        const int xcd = b & 7, j = b >> 3;
        qb = j % a.nqb;
        slice = (j / a.nqb) * 8 + xcd;
    } else {
        slice = b / a.nqb;
        qb = b % a.nqb;
    }
    const int qbase = (qb * waves + wave) * kQueriesPerWave;
    if (qbase >= a.Q) return;  // whole wave is past the last query
    const int q = min(qbase + col, a.Q - 1);

    const int kp = a.kp;
    float *hv = reinterpret_cast<float *>(smem) + (size_t)wave * kQueriesPerWave * kp;
    int *hi = reinterpret_cast<int *>(smem + (size_t)waves * kQueriesPerWave * kp * sizeof(float)) +
              (size_t)wave * kQueriesPerWave * kp;
    for (int e = lane; e < kQueriesPerWave * kp; e += kWave) {
        hv[e] = INFINITY;
        hi[e] = -1;
    }
    wave_sync();
    float *myv = hv + col * kp;
    int *myi = hi + col * kp;
    float theta = INFINITY;

    const long long ldr = a.ldr, ldq = a.ldq;
    const float *__restrict__ qcol = a.query + (long long)h * ldq + q;
    float bq[KS];
    if (!MULTI) {
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) bq[kk] = -2.0f * qcol[(long long)(2 * kk) * ldq];
    }
    const int nchunks = MULTI ? a.nchunks : 1;

    const int r_begin = slice * a.slice_len;
    const int r_end = min(a.R, r_begin + a.slice_len);
    for (int r0 = r_begin; r0 < r_end; r0 += kRefsPerStep) {
        f32x16 acc[kMTiles];
        unsigned voff[kMTiles];
#pragma unroll
        for (int mt = 0; mt < kMTiles; ++mt) {
            const int r = r0 + kRefsPerTile * mt + col;
            const bool ok = r < a.R;
            const int rc = ok ? r : a.R - 1;
            voff[mt] = (unsigned)((long long)h * ldr + rc);
            const float nv = ok ? a.rnorm[rc] : INFINITY;
            const float av = h ? 0.0f : nv;
            const float bv = h ? 0.0f : 1.0f;
            f32x16 z = {};
            acc[mt] = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, z, 0, 0, 0);
        }
        for (int c = 0; c < nchunks; ++c) {
            if (MULTI) {
#pragma unroll
                for (int kk = 0; kk < KS; ++kk)
                    bq[kk] = -2.0f * qcol[(long long)(c * 2 * KS + 2 * kk) * ldq];
            }
            const float *__restrict__ cbase = a.ref + (long long)c * 2 * KS * ldr;
            float abuf[2][CH][kMTiles];
#pragma unroll
            for (int kk = 0; kk < CH; ++kk) {
                const float *__restrict__ row = cbase + (long long)(2 * kk) * ldr;
#pragma unroll
                for (int mt = 0; mt < kMTiles; ++mt) abuf[0][kk][mt] = row[voff[mt]];
            }
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                if (g + 1 < NG) {
#pragma unroll
                    for (int kk = 0; kk < CH; ++kk) {
                        const float *__restrict__ row =
                            cbase + (long long)(2 * ((g + 1) * CH + kk)) * ldr;
#pragma unroll
                        for (int mt = 0; mt < kMTiles; ++mt)
                            abuf[(g + 1) & 1][kk][mt] = row[voff[mt]];
                    }
                }
#pragma unroll
                for (int kk = 0; kk < CH; ++kk) {
#pragma unroll
                    for (int mt = 0; mt < kMTiles; ++mt)
                        acc[mt] = __builtin_amdgcn_mfma_f32_32x32x2f32(
                            abuf[g & 1][kk][mt], bq[g * CH + kk], acc[mt], 0, 0, 0);
                }
            }
        }

        // Epilogue: one compare per key; the slow path runs only when some
        // lane of the wave holds a key below its query's threshold.
        bool cand = false;
#pragma unroll
        for (int mt = 0; mt < kMTiles; ++mt)
#pragma unroll
            for (int i = 0; i < 16; ++i) cand |= acc[mt][i] < theta;
        if (__ballot(cand)) {
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                if (h == hh) {
#pragma unroll
                    for (int mt = 0; mt < kMTiles; ++mt) {
#pragma unroll
                        for (int i = 0; i < 16; ++i) {
                            const float v = acc[mt][i];
                            if (v < theta) {
                                const int r = r0 + kRefsPerTile * mt + acc_row(i, hh);
                                if (r < r_end) theta = heap_replace_top(myv, myi, kp, v, r);
                            }
                        }
                    }
                }
                wave_sync();
                theta = __shfl(theta, col + 32 * hh);
            }
        }
    }

    // Heaps of the wave's 32 consecutive queries are one contiguous block in
    // LDS and in the [slice][query][kp] output.
    wave_sync();
    const long long obase = ((long long)slice * a.Q + qbase) * kp;
    const int nvalid = min(kQueriesPerWave, a.Q - qbase) * kp;
    for (int e = lane; e < nvalid; e += kWave) {
        a.part_val[obase + e] = hv[e];
        a.part_idx[obase + e] = hi[e];
    }
}

// Dimension k-steps with an instantiated scan (padded dimension = 2*KS).
static const int kScanKS[] = {8, 16, 24, 32, 48, 64, 96, 128};

int scan_pick_ks(int dim) {
    for (int ks : kScanKS)
        if (2 * ks >= dim) return ks;
    return 0;  // dimension needs the chunked scan
}

template <int KS, bool MULTI>
static hipError_t launch_scan_t(const ScanArgs &a, int waves, size_t lds, hipStream_t s) {
    auto fn = scan_f32_kernel<KS, MULTI>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)fn,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const int blocks = a.nqb * a.nslices;
    hipLaunchKernelGGL(fn, dim3(blocks), dim3(waves * kWave), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_scan(const ScanArgs &a, int ks, bool multi, int waves, hipStream_t s) {
    const size_t lds = (size_t)waves * kQueriesPerWave * a.kp * (sizeof(float) + sizeof(int));
    if (multi) {
        if (ks != 64) return hipErrorInvalidValue;
        return launch_scan_t<64, true>(a, waves, lds, s);
    }
    switch (ks) {
        case 8: return launch_scan_t<8, false>(a, waves, lds, s);
        case 16: return launch_scan_t<16, false>(a, waves, lds, s);
        case 24: return launch_scan_t<24, false>(a, waves, lds, s);
        case 32: return launch_scan_t<32, false>(a, waves, lds, s);
        case 48: return launch_scan_t<48, false>(a, waves, lds, s);
        case 64: return launch_scan_t<64, false>(a, waves, lds, s);
        case 96: return launch_scan_t<96, false>(a, waves, lds, s);
        case 128: return launch_scan_t<128, false>(a, waves, lds, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace knnhip
