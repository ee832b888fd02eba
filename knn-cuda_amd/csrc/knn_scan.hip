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

__host__ __device__ constexpr int heap_stride(int kp) { return kp | 1; }

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

    // Heaps in LDS with an odd row stride, so the 32 lanes of a half walking
    // their own heaps in step hit distinct banks.
    const int kp = a.kp;
    const int kps = heap_stride(kp);
    float *hv = reinterpret_cast<float *>(smem) + (size_t)wave * kQueriesPerWave * kps;
    int *hi = reinterpret_cast<int *>(smem + (size_t)waves * kQueriesPerWave * kps * sizeof(float)) +
              (size_t)wave * kQueriesPerWave * kps;
    for (int e = lane; e < kQueriesPerWave * kps; e += kWave) {
        hv[e] = INFINITY;
        hi[e] = -1;
    }
    wave_sync();
    float *myv = hv + col * kps;
    int *myi = hi + col * kps;
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

    // Reference operands come through a raw buffer resource per dimension
    // chunk: the lane's byte offset (point + lane-half row) stays in a VGPR,
    // the k-step row offset is a scalar, so a load is one instruction.
    const unsigned kstep_bytes = (unsigned)(2 * ldr * sizeof(float));
    const unsigned span_bytes = (unsigned)min((long long)2 * KS * ldr * (long long)sizeof(float),
                                              (long long)0xFFFFFFFFll);
    auto lane_offsets = [&](int rs, unsigned *vo) {
#pragma unroll
        for (int mt = 0; mt < kMTiles; ++mt) {
            const int r = min(rs + kRefsPerTile * mt + col, a.R - 1);
            vo[mt] = (unsigned)(((long long)h * ldr + r) * (long long)sizeof(float));
        }
    };
    auto chunk_rsrc = [&](int c) {
        return __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float *>(a.ref + (long long)c * 2 * KS * ldr), (short)0, (int)span_bytes,
            0x00020000);
    };
    float abuf[2][CH][kMTiles];
    auto issue = [&](float (&buf)[CH][kMTiles], __amdgpu_buffer_rsrc_t rs, int g,
                     const unsigned *vo) {
#pragma unroll
        for (int kk = 0; kk < CH; ++kk) {
#pragma unroll
            for (int mt = 0; mt < kMTiles; ++mt)
                buf[kk][mt] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                    rs, (int)vo[mt], (int)((g * CH + kk) * kstep_bytes), 0));
        }
    };
    // With an even number of prefetch groups the next step's first group can
    // be fetched while the current step's last group computes.
    constexpr bool kCrossStep = !MULTI && (NG % 2 == 0);

    unsigned voff[kMTiles];
    lane_offsets(r_begin, voff);
    if (kCrossStep && r_begin < r_end) issue(abuf[0], chunk_rsrc(0), 0, voff);

    for (int r0 = r_begin; r0 < r_end; r0 += kRefsPerStep) {
        f32x16 acc[kMTiles];
        float nrm[kMTiles];
#pragma unroll
        for (int mt = 0; mt < kMTiles; ++mt) {
            acc[mt] = f32x16{};
            const int r = r0 + kRefsPerTile * mt + col;
            nrm[mt] = r < a.R ? a.rnorm[r] : INFINITY;
        }
        const int r_next = r0 + kRefsPerStep;
        unsigned voff_next[kMTiles];
        lane_offsets(r_next, voff_next);
        for (int c = 0; c < nchunks; ++c) {
            const __amdgpu_buffer_rsrc_t rs = chunk_rsrc(c);
            if (MULTI) {
#pragma unroll
                for (int kk = 0; kk < KS; ++kk)
                    bq[kk] = -2.0f * qcol[(long long)(c * 2 * KS + 2 * kk) * ldq];
            }
            if (!kCrossStep) issue(abuf[0], rs, 0, voff);
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                if (g + 1 < NG)
                    issue(abuf[(g + 1) & 1], rs, g + 1, voff);
                else if (kCrossStep && r_next < r_end)
                    issue(abuf[0], rs, 0, voff_next);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int kk = 0; kk < CH; ++kk) {
#pragma unroll
                    for (int mt = 0; mt < kMTiles; ++mt)
                        acc[mt] = __builtin_amdgcn_mfma_f32_32x32x2f32(
                            abuf[g & 1][kk][mt], bq[g * CH + kk], acc[mt], 0, 0, 0);
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
#pragma unroll
                for (int mt = 0; mt < kMTiles; ++mt) {
                    bool cm = false;
#pragma unroll
                    for (int i = 0; i < 16; ++i) cm |= acc[mt][i] < theta;
                    if (__ballot(cm && h == hh) == 0) continue;
                    if (h == hh) {
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

    // The wave's 32 consecutive queries are one contiguous block of the
    // [slice][query][kp] output.
    wave_sync();
    const long long obase = ((long long)slice * a.Q + qbase) * kp;
    const int nvalid = min(kQueriesPerWave, a.Q - qbase) * kp;
    for (int e = lane; e < nvalid; e += kWave) {
        const int row = e / kp, j = e - row * kp;
        a.part_val[obase + e] = hv[row * kps + j];
        a.part_idx[obase + e] = hi[row * kps + j];
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

size_t scan_lds_bytes(int kp, int waves) {
    return (size_t)waves * kQueriesPerWave * heap_stride(kp) * (sizeof(float) + sizeof(int));
}

static const void *scan_fn(int ks, bool multi) {
    if (multi) {
        if (ks == 64) return (const void *)scan_f32_kernel<64, true>;
        if (ks == 16) return (const void *)scan_f32_kernel<16, true>;
        return nullptr;
    }
    switch (ks) {
        case 8: return (const void *)scan_f32_kernel<8, false>;
        case 16: return (const void *)scan_f32_kernel<16, false>;
        case 24: return (const void *)scan_f32_kernel<24, false>;
        case 32: return (const void *)scan_f32_kernel<32, false>;
        case 48: return (const void *)scan_f32_kernel<48, false>;
        case 64: return (const void *)scan_f32_kernel<64, false>;
        case 96: return (const void *)scan_f32_kernel<96, false>;
        case 128: return (const void *)scan_f32_kernel<128, false>;
        default: return nullptr;
    }
}

// Resident blocks per CU of one scan instance (VGPR and LDS limits).
int scan_blocks_per_cu(int ks, bool multi, int waves, int kp) {
    const void *fn = scan_fn(ks, multi);
    if (!fn) return 0;
    const size_t lds = scan_lds_bytes(kp, waves);
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

hipError_t launch_scan(const ScanArgs &a, int ks, bool multi, int waves, hipStream_t s) {
    const size_t lds = scan_lds_bytes(a.kp, waves);
    if (multi) {
        if (ks == 64) return launch_scan_t<64, true>(a, waves, lds, s);
        if (ks == 16) return launch_scan_t<16, true>(a, waves, lds, s);
        return hipErrorInvalidValue;
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
