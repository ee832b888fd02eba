// Searches larger than the device memory they may use: the reference set
// streams through two chunk buffers and the queries go in groups, each
// (query group, reference chunk) pair searched exactly on the device and the
// per-chunk results merged there.
//
// Generalises the legacy memory-adaptive split
// (legacy/KNN.CUDA.with.indexes/knn_cuda_with_indexes.cu:340-349 sizes the
// query groups from 90 % of the free memory, :406-444 loops over them),
// which keeps the whole reference set resident and so cannot search a set
// larger than the device.  Here the references are split as well:
//   - a reference chunk of Rc points is searched with the plain engine
//     (run_device), which returns the chunk's exact k nearest in the
//     reference order: ascending (distance, index), distances bit-equal to
//     the ground truth;
//   - the running result and the chunk's are merged by (distance, global
//     index), so after the last chunk each query holds the k smallest pairs of
//     the whole set under the same order, which is the reference's answer;
//   - chunk c's search is queued without waiting for its certificate count,
//     then the host gathers chunk c + 1 into pinned staging and queues its
//     DMA into the other buffer on the upload stream (host_upload) while the
//     device searches chunk c, and only then are chunk c's failed queries,
//     if any, searched again.
#include <algorithm>
#include <cmath>
#include <string>

#include "knn_engine.h"

namespace knnhip {

namespace {

constexpr int kMergeThreads = 256;

// out = the first k of the two sorted lists a (running) and b (one chunk,
// indexes offset by boff), ordered by (distance, index).  One thread per
// query; column q of a k x n result is at [j * n + q].
__global__ __launch_bounds__(kMergeThreads) void merge_topk_kernel(
    const float *__restrict__ ad, const int *__restrict__ ai, const float *__restrict__ bd,
    const int *__restrict__ bi, int boff, int k, int n, float *__restrict__ od,
    int *__restrict__ oi) {
    const int q = blockIdx.x * kMergeThreads + threadIdx.x;
    if (q >= n) return;
    int ia = 0, ib = 0;
    float da = ad[q], db = bd[q];
    int xa = ai[q], xb = bi[q] + boff;
    for (int j = 0; j < k; ++j) {
        const bool take_a = da < db || (da == db && xa < xb);
        od[(size_t)j * n + q] = take_a ? da : db;
        oi[(size_t)j * n + q] = take_a ? xa : xb;
        if (take_a) {
            if (++ia < k) {
                da = ad[(size_t)ia * n + q];
                xa = ai[(size_t)ia * n + q];
            } else {
                da = INFINITY;
                xa = 0x7fffffff;
            }
        } else {
            if (++ib < k) {
                db = bd[(size_t)ib * n + q];
                xb = bi[(size_t)ib * n + q] + boff;
            } else {
                db = INFINITY;
                xb = 0x7fffffff;
            }
        }
    }
}

struct StreamBufs {
    DevBuf ref[2], qry, dist[3], idx[3];
    hipEvent_t loaded[2] = {}, freed[2] = {}, qloaded = {}, qfree = {}, merged = {};
    bool have_events = false;
    Deferred dfr;  // certificate read-back of the chunk being searched
    void release() {
        for (auto &b : ref) b.release();
        qry.release();
        for (auto &b : dist) b.release();
        for (auto &b : idx) b.release();
    }
};
StreamBufs g_sb;

long long g_budget = 0;  // knnhip_set_device_budget; 0: 90 % of free memory

bool grow(DevBuf &b, size_t bytes) { return b.n >= bytes || b.alloc(bytes); }

}  // namespace

size_t search_footprint(int R, int Q, int dim, int k, const knnhip_options &o, bool streamed) {
    const Plan p = make_plan(R, Q, dim, k, o);
    const size_t pm_ld = (size_t)((dim + 3) & ~3);
    size_t b = 0;
    b += (streamed ? 2 : 1) * (size_t)dim * R * 4;      // reference input (two chunk buffers)
    b += (size_t)dim * Q * 4;                            // query input
    b += (streamed ? 3 : 1) * (size_t)k * Q * 8;         // results (chunk + two running)
    if (p.use_scan) {
        b += (size_t)p.r_pad * 4 + (size_t)R * pm_ld * 4;  // norms, point-major rows
        b += (size_t)p.r_pad * p.dpad * 4 * (p.bf3 ? 2 : 1);  // tiles (+ FP32 tiles of rescans)
        b += (size_t)Q * (12 + pm_ld * 4);                 // norms, thresholds, failures, rows
        b += (size_t)p.nslices * p.nqb * block_queries(p) * p.kp * 8;  // partial heaps
        b += (size_t)kMaxEscalation * Q * (4 + (size_t)dim * 4 + (size_t)k * 8);
    }
    b += (size_t)std::min(kExactGrid, Q) * kWave * k * 8;  // exact-scan scratch
    b += (size_t)kNumSlots << 20;                           // 1 MiB rounding of each buffer
    return b;
}

// Largest chunks that fit `budget`: queries in groups of at most 2^17,
// halved while the reference chunk would be small; reference chunks in
// multiples of 1024 points.  False when nothing fits.
bool stream_chunks(int R, int Q, int dim, int k, const knnhip_options &o, size_t budget, int *rc_out,
                   int *qc_out) {
    const int kMinRefChunk = 8192;
    for (int qc = std::min(Q, 1 << 17);; qc = (qc + 1) / 2) {
        int rc = 0;
        if (search_footprint(R, qc, dim, k, o, true) <= budget) {
            rc = R;
        } else {
            int lo = 0, hi = R / 1024;  // in units of 1024 points
            while (lo < hi) {
                const int mid = (lo + hi + 1) / 2;
                if (search_footprint(mid * 1024, qc, dim, k, o, true) <= budget)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            rc = lo * 1024;
        }
        const int nr = rc > 0 ? (int)(((long long)R + rc - 1) / rc) : 0;
        const bool fits = rc > 0 && R / nr >= k;
        if (fits && (rc >= std::min(R, kMinRefChunk) || qc <= 64)) {
            *rc_out = rc;
            *qc_out = qc;
            return true;
        }
        if (qc <= 64) {
            if (rc > 0 && R / nr < k)
                set_error("device memory budget of %zu bytes leaves reference chunks smaller than k=%d",
                          budget, k);
            else
                set_error("device memory budget of %zu bytes is too small for dim=%d k=%d", budget, dim,
                          k);
            return false;
        }
    }
}

// Device memory held in this library's caches (workspace and chunk buffers).
size_t cached_bytes() {
    size_t cached = 0;
    for (size_t c : g_ws.cap) cached += c;
    for (auto *b : {&g_sb.ref[0], &g_sb.ref[1], &g_sb.qry}) cached += b->n;
    for (int j = 0; j < 3; ++j) cached += g_sb.dist[j].n + g_sb.idx[j].n;
    return cached;
}

// Device memory a host-pointer search may use: the caller's budget, else
// 90 % of what is free plus what this library holds in its caches.
size_t device_budget(bool *explicit_budget) {
    *explicit_budget = g_budget > 0;
    if (g_budget > 0) return (size_t)g_budget;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        (void)hipGetLastError();
        return (size_t)-1;
    }
    return (size_t)(0.9 * (double)(free_b + cached_bytes()));
}

void ws_release() {
    (void)hipDeviceSynchronize();
    g_last_fb = nullptr;
    g_last_nonfinite = nullptr;
    for (int s = 0; s < kNumSlots; ++s) {
        if (g_ws.ptr[s]) (void)hipFree(g_ws.ptr[s]);
        g_ws.ptr[s] = nullptr;
        g_ws.cap[s] = 0;
    }
}

bool run_streamed(const float *ref, int R, const float *query, int Q, int dim, int k,
                  float *knn_dist, int *knn_index, const knnhip_options &o, size_t budget,
                  bool kth_only) {
    // buffers cached by earlier calls stay while they fit the budget
    if (cached_bytes() > budget) {
        ws_release();
        g_sb.release();
    }
    // one non-finite flag for every chunk search of the call (the caller
    // checks it at the end)
    if (!nonfinite_clear(nullptr)) return false;
    NonfiniteHold hold;
    int Rc = 0, Qc = 0;
    if (!stream_chunks(R, Q, dim, k, o, budget, &Rc, &Qc)) return false;
    const int nr = (int)(((long long)R + Rc - 1) / Rc);
    const int nq = (int)(((long long)Q + Qc - 1) / Qc);
    auto r_at = [&](int c) { return (int)((long long)R * c / nr); };
    auto q_at = [&](int i) { return (int)((long long)Q * i / nq); };
    {
        char buf[256];
        snprintf(buf, sizeof(buf),
                 "{\"streamed\": true, \"device_budget\": %zu, \"ref_chunk\": %d, \"ref_chunks\": %d, "
                 "\"query_chunk\": %d, \"query_chunks\": %d, \"chunk_plan\": ",
                 budget, Rc, nr, Qc, nq);
        g_last_plan = buf + plan_json(make_plan(r_at(1) - r_at(0), q_at(1) - q_at(0), dim, k, o),
                                      r_at(1) - r_at(0), q_at(1) - q_at(0), dim, k) + "}";
    }
    if (!grow(g_sb.ref[0], (size_t)dim * Rc * 4) || (nr > 1 && !grow(g_sb.ref[1], (size_t)dim * Rc * 4)) ||
        !grow(g_sb.qry, (size_t)dim * Qc * 4)) {
        set_error("Memory allocation error");
        return false;
    }
    for (int j = 0; j < (nr > 1 ? 3 : 1); ++j)
        if (!grow(g_sb.dist[j], (size_t)k * Qc * 4) || !grow(g_sb.idx[j], (size_t)k * Qc * 4)) {
            set_error("Memory allocation error");
            return false;
        }
    if (!g_ws.copy) {
        KNN_TRY(hipStreamCreateWithFlags(&g_ws.copy, hipStreamNonBlocking), "stream create");
        for (auto &e : g_ws.up_ev) KNN_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event create");
        KNN_TRY(hipEventCreateWithFlags(&g_ws.qry_ev, hipEventDisableTiming), "event create");
    }
    if (!g_sb.have_events) {
        for (int b = 0; b < 2; ++b) {
            KNN_TRY(hipEventCreateWithFlags(&g_sb.loaded[b], hipEventDisableTiming), "event create");
            KNN_TRY(hipEventCreateWithFlags(&g_sb.freed[b], hipEventDisableTiming), "event create");
        }
        for (hipEvent_t *e : {&g_sb.qloaded, &g_sb.qfree, &g_sb.merged})
            KNN_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming), "event create");
        g_sb.have_events = true;
    }
    hipStream_t s = nullptr, cs = g_ws.copy;
    for (int b = 0; b < 2; ++b) KNN_TRY(hipEventRecord(g_sb.freed[b], s), "event");

    // Reference chunk c into buffer b on the upload stream, after the search
    // that last read the buffer.
    auto copy_chunk = [&](int c, int b) -> bool {
        const int r0 = r_at(c), r1 = r_at(c + 1);
        KNN_TRY(hipStreamWaitEvent(cs, g_sb.freed[b], 0), "event wait");
        // the chunk's rows packed at leading dimension r1 - r0
        if (!host_upload(g_sb.ref[b].p, r1 - r0, ref + r0, R, r1 - r0, dim, cs)) return false;
        KNN_TRY(hipEventRecord(g_sb.loaded[b], cs), "event");
        return true;
    };
    // No return while a copy may still read the caller's buffer.
    struct Drain {
        hipStream_t cs;
        ~Drain() { (void)hipStreamSynchronize(cs); }
    } drain{cs};

    // Every copy between host and device goes on the upload stream, through
    // pinned staging, in program order; events order the copies against the
    // searches.
    if (!copy_chunk(0, 0)) return false;
    const int total = nq * nr;
    int cur = 0;  // running result buffer
    for (int it = 0; it < total; ++it) {
        const int qi = it / nr, c = it % nr;
        const int q0 = q_at(qi), nqc = q_at(qi + 1) - q0;
        const int r0 = r_at(c), nrc = r_at(c + 1) - r0;
        const int rb = nr > 1 ? it % 2 : 0;
        if (c == 0) {
            // the group's queries, once the previous group's searches are done with the buffer
            KNN_TRY(hipEventRecord(g_sb.qfree, s), "event");
            KNN_TRY(hipStreamWaitEvent(cs, g_sb.qfree, 0), "event wait");
            if (!host_upload(g_sb.qry.p, nqc, query + q0, Q, nqc, dim, cs)) return false;
            KNN_TRY(hipEventRecord(g_sb.qloaded, cs), "event");
            KNN_TRY(hipStreamWaitEvent(s, g_sb.qloaded, 0), "event wait");
            cur = 0;
        }
        KNN_TRY(hipStreamWaitEvent(s, g_sb.loaded[rb], 0), "event wait");
        const int out = c == 0 ? cur : 2;
        const Plan p = make_plan(nrc, nqc, dim, k, o);
        // queued without waiting for the certificate count; the next chunk's
        // copy is issued while this chunk is searched, then the failed
        // queries (if any) are searched again
        if (!run_device(g_sb.ref[rb].as<float>(), nrc, nrc, g_sb.qry.as<float>(), nqc, nqc, dim, k,
                        g_sb.dist[out].as<float>(), g_sb.idx[out].as<int>(), p, o, s, 0, nullptr,
                        &g_sb.dfr))
            return false;
        if (nr > 1 && it + 1 < total && !copy_chunk((it + 1) % nr, rb ^ 1)) return false;
        if (!finish_device(g_sb.dfr)) return false;
        if (nr > 1) KNN_TRY(hipEventRecord(g_sb.freed[rb], s), "event");
        if (c > 0) {
            hipLaunchKernelGGL(merge_topk_kernel, dim3((nqc + kMergeThreads - 1) / kMergeThreads),
                               dim3(kMergeThreads), 0, s, g_sb.dist[cur].as<float>(),
                               g_sb.idx[cur].as<int>(), g_sb.dist[2].as<float>(), g_sb.idx[2].as<int>(),
                               r0, k, nqc, g_sb.dist[cur ^ 1].as<float>(), g_sb.idx[cur ^ 1].as<int>());
            KNN_TRY(hipGetLastError(), "merge kernel launch");
            cur ^= 1;
        }
        if (c == nr - 1) {
            // the group's results, behind its last merge
            KNN_TRY(hipEventRecord(g_sb.merged, s), "event");
            KNN_TRY(hipStreamWaitEvent(cs, g_sb.merged, 0), "event wait");
            const float *dd = g_sb.dist[cur].as<float>();
            const int *ii = g_sb.idx[cur].as<int>();
            if (kth_only) {
                if (!host_download(knn_dist + q0, Q, dd + (size_t)(k - 1) * nqc, nqc, nqc, 1, cs))
                    return false;
            } else if (!host_download(knn_dist + q0, Q, dd, nqc, nqc, k, cs) ||
                       !host_download(knn_index + q0, Q, ii, nqc, nqc, k, cs)) {
                return false;
            }
            // the next group overwrites the result buffers
            KNN_TRY(hipEventRecord(g_sb.qfree, cs), "event");
            KNN_TRY(hipStreamWaitEvent(s, g_sb.qfree, 0), "event wait");
        }
    }
    KNN_TRY(hipStreamSynchronize(cs), "Unable to copy data from device to host");
    KNN_TRY(hipStreamSynchronize(s), "Unable to execute kernel");
    return true;
}

}  // namespace knnhip

using namespace knnhip;

extern "C" {

int knnhip_set_device_budget(long long bytes) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (bytes < 0) {
        set_error("device budget must be >= 0 (got %lld)", bytes);
        return 0;
    }
    g_budget = bytes;
    return 1;
}

long long knnhip_device_budget(void) {
    std::lock_guard<std::mutex> lock(g_mu);
    return g_budget;
}

int knnhip_stream_chunks(int ref_nb, int query_nb, int dim, int k, long long device_bytes,
                         int *ref_chunk, int *query_chunk) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!ref_chunk || !query_chunk) {
        set_error("null pointer argument");
        return 0;
    }
    if (ref_nb < 1 || query_nb < 1 || dim < 1 || k < 1 || k > ref_nb || device_bytes <= 0) {
        set_error("invalid sizes (ref_nb=%d query_nb=%d dim=%d k=%d device_bytes=%lld)", ref_nb,
                  query_nb, dim, k, device_bytes);
        return 0;
    }
    if (!ensure_device()) return 0;
    return stream_chunks(ref_nb, query_nb, dim, k, default_options(0), (size_t)device_bytes, ref_chunk,
                         query_chunk)
               ? 1
               : 0;
}

}  // extern "C"
