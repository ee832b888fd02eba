// Host side of the plain exact k-NN engine: argument checks, launch
// planning, device workspace, and the drop-in entry points.
//
// Replaces the host orchestration of code/knncuda.cu:271-415
// (knn_cuda_global), :418-609 (knn_cuda_texture) and :612-832 (knn_cublas):
// device 0, host pointers in and out, synchronous, false + an "ERROR:" line
// on stdout on failure.  Unlike the reference, device buffers are cached
// between calls (grown on demand, never visible to the caller), the R x Q
// distance matrix never exists, and all three entry points return the
// reference ground truth exactly.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "knn_engine.h"
#include "knn_index.h"

namespace knnhip {

std::mutex g_mu;
std::string g_last_error;
float g_stage_ms[4] = {-1.f, -1.f, -1.f, -1.f};
hipStream_t g_last_stream = nullptr;
const int *g_last_fb = nullptr;  // device counter of the last search's exact-scan queries
std::string g_last_plan;
int g_cu_count = 0;
Workspace g_ws;

void set_error(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

Arena *g_arena = nullptr;

void *ws_get(Slot s, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (g_arena) {
        // a slot asked for again within one call gets its piece back
        Arena &a = *g_arena;
        if (a.slot[s] && a.size[s] >= bytes) return a.slot[s];
        const size_t need = (bytes + 255) & ~(size_t)255;
        if (a.used + need > a.cap) return nullptr;
        a.slot[s] = a.base + a.used;
        a.size[s] = need;
        a.used += need;
        return a.slot[s];
    }
    if (g_ws.cap[s] >= bytes) return g_ws.ptr[s];
    if (g_ws.ptr[s]) {
        (void)hipFree(g_ws.ptr[s]);
        g_ws.ptr[s] = nullptr;
        g_ws.cap[s] = 0;
    }
    void *p = nullptr;
    const size_t rounded = (bytes + 0xFFFFF) & ~(size_t)0xFFFFF;
    if (hipMalloc(&p, rounded) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    g_ws.ptr[s] = p;
    g_ws.cap[s] = rounded;
    return p;
}

const unsigned *g_last_nonfinite = nullptr;
namespace {
int g_nonfinite_holds = 0;
}  // namespace
NonfiniteHold::NonfiniteHold() { ++g_nonfinite_holds; }
NonfiniteHold::~NonfiniteHold() { --g_nonfinite_holds; }

bool nonfinite_clear(hipStream_t s) {
    unsigned *f = (unsigned *)ws_get(kNonFinite, sizeof(unsigned));
    if (!f) {
        set_error("Memory allocation error");
        return false;
    }
    KNN_TRY(hipMemsetAsync(f, 0, sizeof(unsigned), s), "memset");
    return true;
}

bool nonfinite_check(hipStream_t s) {
    const unsigned *f = (const unsigned *)ws_get(kNonFinite, sizeof(unsigned));
    unsigned bad = 0;
    if (!f) {
        set_error("Memory allocation error");
        return false;
    }
    KNN_TRY(hipMemcpyAsync(&bad, f, sizeof(unsigned), hipMemcpyDeviceToHost, s), "copy");
    KNN_TRY(hipStreamSynchronize(s), "sync");
    if (bad) {
        set_error("non-finite input: a reference or query point has a NaN or infinite squared norm");
        return false;
    }
    return true;
}

namespace {

// Cost model of one slice count: the scan runs in `rounds` waves of
// resident blocks, each block's `waves` waves sharing ceil(steps / S)
// reference steps, plus a warm-up of about two steps while the heaps fill.
double slice_cost(int S, int steps, int nqb, int resident, int waves) {
    const long long units = (long long)nqb * S;
    const long long rounds = (units + resident - 1) / resident;
    return (double)rounds * ((double)((steps + S - 1) / S) / waves + 2.0);
}

}  // namespace

Plan make_plan(int R, int Q, int dim, int k, const knnhip_options &o) {
    Plan p;
    const int margin = o.margin > 0 ? o.margin : default_margin(k);
    p.kp = std::min(k + margin, R);
    if (p.kp < k) p.kp = k;
    // Heaps: 32 queries * (kp | 1) * 8 bytes of LDS per block, shared by its
    // waves.
    if (scan_lds_bytes(p.kp, 4) > kScanLdsMax) return p;  // k too large for the scan's heaps
    p.waves = 4;
    p.ks = scan_pick_ks(dim);
    if (p.ks) {
        p.nchunks = 1;
        p.bmode = kBRegs;
        p.dpad = 2 * p.ks;
    } else if (scan_block_lds(p.kp, 4, kBLds, (dim + 127) / 128 * 16) <= kScanLdsMax) {
        // wider dimensions: the block's query rows staged in LDS.  Padded to
        // 128 dimensions like the chunked layout, so that a search with a
        // larger kp over the same packed tiles can take the chunked instance.
        p.ks = 16;
        p.bmode = kBLds;
        p.nchunks = (dim + 127) / 128;
        p.dpad = 128 * p.nchunks;
    } else {
        // ... or, where they do not fit beside the heaps, reloaded into VGPRs
        // 128 dimensions at a time
        p.ks = 64;
        p.bmode = kBChunks;
        p.nchunks = (dim + 2 * p.ks - 1) / (2 * p.ks);
        p.dpad = 2 * p.ks * p.nchunks;
    }
    p.oct = p.dpad / 8;
    p.nqb = (Q + kQueriesPerWave - 1) / kQueriesPerWave;
    // Slices: the count whose blocks fill whole rounds of resident blocks
    // (cost model above), bounded by the merge capacity of the select stage
    // and by one step per slice.
    const int steps = (R + kRefsPerStep - 1) / kRefsPerStep;
    p.r_pad = steps * kRefsPerStep;
    // A slice's packed tiles are addressed with 32-bit offsets.
    const long long step_bytes = (long long)kRefsPerStep * p.dpad * sizeof(float);
    const long long max_steps_per_slice = 0xFFFFFFFFll / step_bytes;
    const int smin = (int)((steps + max_steps_per_slice - 1) / max_steps_per_slice);
    const int smax = std::max(1, std::min(kMaxUnion / p.kp, steps));
    if (smin > smax) return p;  // too large for one scan launch
    if (g_cu_count <= 0) {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
            g_cu_count = n;
        else
            (void)hipGetLastError();
    }
    p.cus = g_cu_count > 0 ? g_cu_count : 256;
    p.blocks_per_cu = std::max(1, scan_blocks_per_cu(p.ks, p.bmode, p.waves, p.kp, p.oct));
    if (p.blocks_per_cu == 1 && scan_block_lds(p.kp, 8, p.bmode, p.oct) <= kScanLdsMax) {
        // LDS leaves room for one block per CU: eight waves on its queries
        // keep two waves per SIMD
        p.waves = 8;
        p.blocks_per_cu = std::max(1, scan_blocks_per_cu(p.ks, p.bmode, p.waves, p.kp, p.oct));
    }
    const int resident = p.cus * p.blocks_per_cu;
    int s = smin;
    if (o.slices > 0) {
        s = std::max(smin, std::min(o.slices, smax));
    } else {
        double best = slice_cost(smin, steps, p.nqb, resident, p.waves);
        for (int c = smin + 1; c <= smax; ++c) {
            const double cost = slice_cost(c, steps, p.nqb, resident, p.waves);
            if (cost < best) {
                best = cost;
                s = c;
            }
        }
    }
    const int steps_per_slice = (steps + s - 1) / s;
    p.slice_len = steps_per_slice * kRefsPerStep;
    p.slice_span = (long long)steps_per_slice * step_bytes;
    p.nslices = s;
    p.xcd = (s % 8 == 0) ? 1 : 0;
    p.maxc = std::min(kMaxCand, std::max(2 * p.kp, 256));
    p.use_scan = true;
    if (o.precision >= 1 || o.block_queries == 128) {
        // LDS-shared scan: 128 queries per block, one block per CU; slices
        // fill whole rounds of CUs (the same cost model, resident = CUs).
        const int nch = scan_bf3_nch(dim);
        if (nch && scan_bf3_lds_bytes(p.kp) <= kScanLdsMax) {
            p.lds = true;
            p.bf3 = o.precision >= 1;
            p.nch = nch;
            p.waves = 8;
            p.dpad = 16 * nch;
            // octets written by the norms pass: the dimension's (bfloat16
            // tiles zero-fill the rest), or every octet of the FP32 tiles
            p.oct = p.bf3 ? (dim + 7) / 8 : 2 * nch;
            p.nqb = (Q + 127) / 128;
            p.blocks_per_cu = 1;
            const long long bstep = (long long)kRefsPerStep * p.dpad * sizeof(float);
            const int bmin = (int)((steps + 0x7FFFFFFFll / bstep - 1) / (0x7FFFFFFFll / bstep));
            int bs = std::max(1, bmin);
            if (o.slices > 0) {
                bs = std::max(bmin, std::min(o.slices, smax));
            } else {
                double best = slice_cost(bs, steps, p.nqb, p.cus, 4);
                for (int c = bs + 1; c <= smax; ++c) {
                    const double cost = slice_cost(c, steps, p.nqb, p.cus, 4);
                    if (cost < best) {
                        best = cost;
                        bs = c;
                    }
                }
            }
            const int sps = (steps + bs - 1) / bs;
            p.slice_len = sps * kRefsPerStep;
            p.slice_span = (long long)sps * bstep;
            p.nslices = (steps + sps - 1) / sps;
            p.xcd = 0;
        }
    }
    return p;
}

std::string plan_json(const Plan &p, int R, int Q, int dim, int k) {
    char buf[512];
    snprintf(buf, sizeof(buf),
             "{\"ref_nb\": %d, \"query_nb\": %d, \"dim\": %d, \"k\": %d, \"scan\": %s, "
             "\"ks\": %d, \"b_operands\": \"%s\", \"chunks\": %d, \"dpad\": %d, \"waves_per_block\": %d, \"kp\": %d, "
             "\"slices\": %d, \"slice_len\": %d, \"query_blocks\": %d, \"xcd_map\": %d, "
             "\"blocks_per_cu\": %d, \"cus\": %d, \"max_candidates\": %d, \"precision\": \"%s\", \"block_queries\": %d}",
             R, Q, dim, k, p.use_scan ? "true" : "false", p.ks,
             p.bmode == kBLds ? "lds" : p.bmode == kBChunks ? "chunks" : "registers", p.nchunks, p.dpad,
             p.waves, p.kp,
             p.nslices, p.slice_len, p.nqb, p.xcd, p.blocks_per_cu, p.cus, p.maxc,
             p.bf3 ? "bf16x3" : "fp32", p.lds ? 128 : 32);
    return buf;
}

bool check_args(const void *ref, int R, const void *query, int Q, int dim, int k, const void *dist,
                const void *idx) {
    if (!ref || !query || !dist || !idx) {
        set_error("null pointer argument");
        return false;
    }
    if (R < 1 || Q < 1 || dim < 1) {
        set_error("sizes must be positive (ref_nb=%d query_nb=%d dim=%d)", R, Q, dim);
        return false;
    }
    if (k < 1 || k > R) {
        set_error("k must satisfy 1 <= k <= ref_nb (k=%d ref_nb=%d)", k, R);
        return false;
    }
    return true;
}

bool ensure_device() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        set_error("No HIP device found");
        return false;
    }
    if (hipSetDevice(0) != hipSuccess) {
        (void)hipGetLastError();
        set_error("Cannot set the HIP device");
        return false;
    }
    return true;
}

namespace {
hipEvent_t g_call_ev = nullptr;     // end of the last call's work
hipStream_t g_call_stream = nullptr;
bool g_call_recorded = false;
}  // namespace

CallOrder::CallOrder(hipStream_t s) : s_(s) {
    if (g_call_recorded && g_call_stream != s && hipStreamWaitEvent(s, g_call_ev, 0) != hipSuccess) {
        // cannot order on the device: wait on the host instead
        (void)hipGetLastError();
        (void)hipEventSynchronize(g_call_ev);
    }
}

CallOrder::~CallOrder() {
    if (!g_call_ev && hipEventCreateWithFlags(&g_call_ev, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        g_call_ev = nullptr;
        return;
    }
    if (hipEventRecord(g_call_ev, s_) != hipSuccess) {
        (void)hipGetLastError();
        g_call_recorded = false;
        return;
    }
    g_call_stream = s_;
    g_call_recorded = true;
}

// Key error bound of the precision scan (see knn_scan_bf16.hip): with
// u = 2^-8, n = 3 * 16 * nch + 1 FP32 additions (unit 2^-23, allowing for any
// rounding inside the MFMA) and the FP32 norm over dim terms,
//   err_q = 3.01 u^2 + 1.02 g_n,  err_r = g_dim + 2.03 g_n + 3.01 u^2,
// widened by 5 %.
void bf3_error_bound(int dim, int nch, double *err_q, double *err_r) {
    const double u2 = 1.0 / 65536.0;
    const double n = 3.0 * 16.0 * nch + 1.0;
    const double gn = n * 1.1920928955078125e-7 / (1.0 - n * 1.1920928955078125e-7);
    const double gd = dim * 5.9604644775390625e-8 / (1.0 - dim * 5.9604644775390625e-8);
    *err_q = 1.05 * (3.01 * u2 + 1.02 * gn);
    *err_r = 1.05 * (gd + 2.03 * gn + 3.01 * u2);
}

namespace {

// Buffers of one escalation level: the failed queries' ids, their columns
// and their results.
struct Escalation {
    DevBuf list, qsub, dsub, isub;
};
Escalation g_esc[kMaxEscalation];

}  // namespace

// The nfail queries of d's search that failed the certificate (ids in
// d.fb_list) are searched again through the scan with margin d.margin and
// FP32 products, and their results scattered into d's outputs.
bool escalate(const Deferred &d, int nfail) {
    if (nfail <= 0) return true;
    Escalation &e = g_esc[d.level];
    auto need = [](DevBuf &b, size_t bytes) { return b.n >= bytes || b.alloc(bytes); };
    if (!need(e.list, (size_t)nfail * 4) || !need(e.qsub, (size_t)nfail * d.dim * 4) ||
        !need(e.dsub, (size_t)nfail * d.k * 4) || !need(e.isub, (size_t)nfail * d.k * 4)) {
        set_error("Memory allocation error");
        return false;
    }
    const hipStream_t s = d.s;
    KNN_TRY(hipMemcpyAsync(e.list.p, d.fb_list, (size_t)nfail * 4, hipMemcpyDeviceToDevice, s), "copy");
    KNN_TRY(idx_gather_columns(d.query, d.ldq, d.dim, e.list.as<int>(), nfail, e.qsub.as<float>(), s),
            "gather");
    knnhip_options o2 = d.o;
    o2.margin = d.margin;
    o2.force_fallback = 0;
    o2.timing = 0;
    o2.precision = 0;  // the rescans use FP32 products
    const Plan p2 = make_plan(d.R, nfail, d.dim, d.k, o2);
    if (!run_device(d.ref, d.ldr, d.R, e.qsub.as<float>(), nfail, nfail, d.dim, d.k, e.dsub.as<float>(),
                    e.isub.as<int>(), p2, o2, s, d.level + 1))
        return false;
    KNN_TRY(idx_scatter_results(e.dsub.as<float>(), e.isub.as<int>(), d.k, nfail, e.list.as<int>(), d.Q,
                                d.out_dist, d.out_idx, s),
            "scatter");
    g_last_fb = d.fb_count;
    return true;
}

bool finish_device(Deferred &d) {
    if (!d.pending) return true;
    d.pending = false;
    KNN_TRY(hipEventSynchronize(d.ev), "event sync");
    return escalate(d, *d.nfail_host);
}

// Core: device pointers, dim rows of ldr / ldq floats.
bool run_device(const float *ref, long long ldr, int R, const float *query, long long ldq, int Q,
                int dim, int k, float *out_dist, int *out_idx, const Plan &p,
                const knnhip_options &o, hipStream_t s, int level, const Upload *up, Deferred *dfr) {
    if (up && (!p.use_scan || p.lds || p.bf3 || up->chunks > kUploadChunks)) {
        set_error("chunked upload needs an FP32 32-query scan plan");
        return false;
    }
    if (dfr) dfr->pending = false;
    const bool timing = o.timing != 0;
    if (timing && !g_ws.have_events) {
        // (stage timestamps only: no system-scope fence)
        for (auto &e : g_ws.ev) KNN_TRY(hipEventCreateWithFlags(&e, hipEventDisableSystemFence), "event create");
        g_ws.have_events = true;
    }
    for (float &m : g_stage_ms) m = -1.f;
    int *fb_count = (int *)ws_get(kFbCount, sizeof(int));
    unsigned *nonfinite = (unsigned *)ws_get(kNonFinite, sizeof(unsigned));
    if (!fb_count || !nonfinite) {
        set_error("Memory allocation error");
        return false;
    }
    KNN_TRY(hipMemsetAsync(fb_count, 0, sizeof(int), s), "memset");
    if (level == 0 && g_nonfinite_holds == 0)
        KNN_TRY(hipMemsetAsync(nonfinite, 0, sizeof(unsigned), s), "memset");
    g_last_nonfinite = nonfinite;

    ExactArgs ea{};
    ea.ref = ref;
    ea.ldr = ldr;
    ea.R = R;
    ea.query = query;
    ea.ldq = ldq;
    ea.Q = Q;
    ea.dim = dim;
    ea.k = k;
    ea.out_dist = out_dist;
    ea.out_idx = out_idx;
    ea.nonfinite = nonfinite;
    const int egrid = std::min(kExactGrid, Q);
    ea.scratch_key = (float *)ws_get(kFbKey, (size_t)egrid * kWave * k * sizeof(float));
    ea.scratch_idx = (int *)ws_get(kFbIdx, (size_t)egrid * kWave * k * sizeof(int));
    if (!ea.scratch_key || !ea.scratch_idx) {
        set_error("Memory allocation error");
        return false;
    }

    if (!p.use_scan) {
        // the norms only for the non-finite check
        float *rn = (float *)ws_get(kRNorm, (size_t)R * sizeof(float));
        float *qn = (float *)ws_get(kQNorm, (size_t)Q * sizeof(float));
        if (!rn || !qn) {
            set_error("Memory allocation error");
            return false;
        }
        for (int side = 0; side < 2; ++side) {
            NormArgs nc{};
            nc.p = side ? query : ref;
            nc.ld = side ? ldq : ldr;
            nc.n = nc.n_pad = side ? Q : R;
            nc.dim = dim;
            nc.oct = (dim + 7) / 8;
            nc.out = side ? qn : rn;
            nc.nonfinite = nonfinite;
            KNN_TRY(launch_norms(nc, s), "norm kernel launch");
        }
        if (timing) KNN_TRY(hipEventRecord(g_ws.ev[3], s), "event");
        KNN_TRY(launch_exact(ea, egrid, s), "exact scan launch");
        if (timing) KNN_TRY(hipEventRecord(g_ws.ev[4], s), "event");
    } else {
        float *rnorm = (float *)ws_get(kRNorm, (size_t)p.r_pad * sizeof(float));
        float *qnorm = (float *)ws_get(kQNorm, (size_t)Q * sizeof(float));
        unsigned *rmax = (unsigned *)ws_get(kRMax, sizeof(unsigned));
        const size_t npart = (size_t)p.nslices * p.nqb * block_queries(p) * p.kp;
        float *pval = (float *)ws_get(kPartVal, npart * sizeof(float));
        int *pidx = (int *)ws_get(kPartIdx, npart * sizeof(int));
        int *fb_list = (int *)ws_get(kFbList, (size_t)Q * sizeof(int));
        const int pm_ld = (dim + 3) & ~3;
        float *ref_pm = (float *)ws_get(kRefPm, (size_t)R * pm_ld * sizeof(float));
        float *qry_pm = (float *)ws_get(kQryPm, (size_t)Q * pm_ld * sizeof(float));
        // FP32 tiles, or bfloat16 hi/lo tiles (the same bytes) for the precision scan
        void *packed = ws_get(p.bf3 ? kPackedBf : kPacked, (size_t)p.r_pad * p.dpad * sizeof(float));
        unsigned *theta = (unsigned *)ws_get(kTheta, (size_t)Q * sizeof(unsigned));
        if (!rnorm || !qnorm || !rmax || !pval || !pidx || !fb_list || !ref_pm || !qry_pm ||
            !packed || !theta) {
            set_error("Memory allocation error");
            return false;
        }
        if (timing) KNN_TRY(hipEventRecord(g_ws.ev[0], s), "event");
        KNN_TRY(hipMemsetAsync(rmax, 0, sizeof(unsigned), s), "memset");
        if (p.lds) KNN_TRY(hipMemsetAsync(rnorm, 0, (size_t)p.r_pad * sizeof(float), s), "memset");
        KNN_TRY(hipMemsetD32Async((hipDeviceptr_t)theta, 0xFF800000u /* +inf, order-mapped */, Q, s),
                "memset");
        NormArgs na{};
        na.p = ref;
        na.ld = ldr;
        na.n = R;
        na.n_pad = p.r_pad;
        na.dim = dim;
        na.oct = p.oct;
        na.out = rnorm;
        na.maxbits = rmax;
        na.pm = ref_pm;
        na.pm_ld = pm_ld;
        na.nonfinite = nonfinite;
        if (p.bf3) {
            na.packed_bf = (unsigned short *)packed;
            na.nch_bf = p.nch;
        } else
            na.packed = (float *)packed;
        if (!up) KNN_TRY(launch_norms(na, s), "norm kernel launch");
        NormArgs nq{};
        nq.p = query;
        nq.ld = ldq;
        nq.n = Q;
        nq.n_pad = Q;
        nq.dim = dim;
        nq.oct = (dim + 7) / 8;
        nq.out = qnorm;
        nq.pm = qry_pm;
        nq.pm_ld = pm_ld;
        nq.nonfinite = nonfinite;
        KNN_TRY(launch_norms(nq, s), "norm kernel launch");
        if (timing) KNN_TRY(hipEventRecord(g_ws.ev[1], s), "event");

        ScanArgs sa{};
        sa.packed = p.bf3 ? nullptr : (const float *)packed;
        sa.packed_bf = p.bf3 ? (const unsigned short *)packed : nullptr;
        sa.oct = p.oct;
        sa.R = R;
        sa.dim = dim;
        sa.theta = theta;
        sa.slice_span = p.slice_span;
        sa.r_pad = p.r_pad;
        sa.query = query;
        sa.ldq = ldq;
        sa.Q = Q;
        sa.rnorm = rnorm;
        sa.slice_len = p.slice_len;
        sa.nslices = p.nslices;
        sa.kp = p.kp;
        sa.nqb = p.nqb;
        sa.nchunks = p.nchunks;
        sa.xcd_map = p.xcd;
        sa.part_val = pval;
        sa.part_idx = pidx;
        if (up) {
            // chunk g: slices [s0, s1), references [r0, r1); its copy lands on
            // the upload stream while the search stream works on chunk g - 1
            const int G = std::min(up->chunks, p.nslices);
            for (int g = 0; g < G; ++g) {
                const int s0 = (int)((long long)p.nslices * g / G);
                const int s1 = (int)((long long)p.nslices * (g + 1) / G);
                const int r0 = s0 * p.slice_len, r1 = std::min(R, s1 * p.slice_len);
                if (s1 <= s0 || r1 <= r0) continue;
                if (!host_upload(const_cast<float *>(ref) + r0, ldr, up->ref + r0, up->ld, r1 - r0, dim,
                                 up->stream))
                    return false;
                KNN_TRY(hipEventRecord(g_ws.up_ev[g], up->stream), "event");
                KNN_TRY(hipStreamWaitEvent(s, g_ws.up_ev[g], 0), "event wait");
                NormArgs nc = na;
                nc.p = ref + r0;
                nc.n = r1 - r0;
                nc.n_pad = (s1 == p.nslices ? p.r_pad : s1 * p.slice_len) - r0;
                nc.out = rnorm + r0;
                nc.pm = ref_pm + (size_t)r0 * pm_ld;
                nc.packed = (float *)packed + (size_t)(r0 / kRefsPerTile) * p.oct * kWave * 4;
                KNN_TRY(launch_norms(nc, s), "norm kernel launch");
                ScanArgs sc = sa;
                sc.slice0 = s0;
                sc.nslices = s1 - s0;
                sc.xcd_map = 0;
                KNN_TRY(launch_scan(sc, p.ks, p.bmode, p.waves, s), "scan kernel launch");
            }
        } else if (p.lds) {
            KNN_TRY(launch_scan_lds(sa, p.nch, p.bf3, s), "scan kernel launch");
        } else {
            KNN_TRY(launch_scan(sa, p.ks, p.bmode, p.waves, s), "scan kernel launch");
        }
        if (timing) KNN_TRY(hipEventRecord(g_ws.ev[2], s), "event");

        FinalArgs fa{};
        fa.ref = ref;
        fa.ldr = ldr;
        fa.R = R;
        fa.query = query;
        fa.ldq = ldq;
        fa.Q = Q;
        fa.dim = dim;
        fa.k = k;
        fa.kp = p.kp;
        fa.nrows = p.nslices;
        fa.row_stride = (long long)p.nqb * block_queries(p);
        if (p.bf3) bf3_error_bound(dim, p.nch, &fa.err_q, &fa.err_r);
        fa.maxc = p.maxc;
        fa.part_val = pval;
        fa.part_idx = pidx;
        fa.ref_pm = ref_pm;
        fa.query_pm = qry_pm;
        fa.pm_ld = pm_ld;
        fa.qnorm = qnorm;
        fa.rmax = rmax;
        fa.theta = theta;
        fa.force_fallback = o.force_fallback;
        fa.out_dist = out_dist;
        fa.out_idx = out_idx;
        fa.fb_count = fb_count;
        fa.fb_list = fb_list;
        fa.nonfinite = nonfinite;
        KNN_TRY(launch_select(fa, s), "select kernel launch");
        if (timing) KNN_TRY(hipEventRecord(g_ws.ev[3], s), "event");

        // Queries that failed the certificate: once more through the scan
        // with a wider margin while one is left, else the exact per-query
        // scan (whose grid reads the count on the device).  A forced
        // fallback goes to the exact scan directly.
        const int cur_margin = p.kp - k;
        int next_margin = 0;
        for (int l = level; l < kMaxEscalation && !next_margin && !o.force_fallback && !g_arena; ++l)
            if (kEscalationMargins[l] > cur_margin && k + cur_margin < R) next_margin = kEscalationMargins[l];
        if (next_margin) {
            Deferred local;
            Deferred &d = dfr ? *dfr : local;
            d.ref = ref;
            d.ldr = ldr;
            d.R = R;
            d.query = query;
            d.ldq = ldq;
            d.Q = Q;
            d.dim = dim;
            d.k = k;
            d.out_dist = out_dist;
            d.out_idx = out_idx;
            d.fb_list = fb_list;
            d.fb_count = fb_count;
            d.o = o;
            d.margin = next_margin;
            d.level = level;
            d.s = s;
            if (dfr) {
                // the caller reads the count back later (finish_device)
                if (!d.nfail_host) {
                    KNN_TRY(hipHostMalloc((void **)&d.nfail_host, sizeof(int), hipHostMallocDefault),
                            "host alloc");
                    KNN_TRY(hipEventCreateWithFlags(&d.ev, hipEventDisableTiming), "event create");
                }
                KNN_TRY(hipMemcpyAsync(d.nfail_host, fb_count, sizeof(int), hipMemcpyDeviceToHost, s),
                        "copy");
                KNN_TRY(hipEventRecord(d.ev, s), "event");
                d.pending = true;
            } else {
                int nfail = 0;
                KNN_TRY(hipMemcpyAsync(&nfail, fb_count, sizeof(int), hipMemcpyDeviceToHost, s), "copy");
                KNN_TRY(hipStreamSynchronize(s), "sync");
                if (!escalate(d, nfail)) return false;
            }
        } else {
            ea.list = fb_list;
            ea.count = fb_count;
            KNN_TRY(launch_exact(ea, egrid, s), "exact scan launch");
        }
        if (timing) KNN_TRY(hipEventRecord(g_ws.ev[4], s), "event");
    }
    g_last_stream = s;
    g_last_fb = fb_count;
    if (timing) {
        KNN_TRY(hipEventSynchronize(g_ws.ev[4]), "event sync");
        if (p.use_scan) {
            for (int i = 0; i < 4; ++i)
                KNN_TRY(hipEventElapsedTime(&g_stage_ms[i], g_ws.ev[i], g_ws.ev[i + 1]),
                        "event time");
        } else {
            KNN_TRY(hipEventElapsedTime(&g_stage_ms[3], g_ws.ev[3], g_ws.ev[4]), "event time");
        }
    }
    return true;
}

size_t run_device_bytes(const Plan &p, int R, int Q, int dim, int k) {
    auto piece = [](size_t b) { return ((b ? b : 256) + 255) & ~(size_t)255; };
    const int egrid = std::min(kExactGrid, Q);
    size_t n = piece(sizeof(int)) + piece(sizeof(unsigned)) + 2 * piece((size_t)egrid * kWave * k * 4);
    if (!p.use_scan) return n + piece((size_t)R * 4) + piece((size_t)Q * 4);
    const int pm_ld = (dim + 3) & ~3;
    const size_t npart = (size_t)p.nslices * p.nqb * block_queries(p) * p.kp;
    n += piece((size_t)p.r_pad * 4) + piece((size_t)Q * 4) + piece(4) + 2 * piece(npart * 4) +
         piece((size_t)Q * 4) + piece((size_t)R * pm_ld * 4) + piece((size_t)Q * pm_ld * 4) +
         piece((size_t)p.r_pad * p.dpad * 4) + piece((size_t)Q * 4);
    return n;
}

knnhip_options default_options(int variant) {
    knnhip_options o;
    memset(&o, 0, sizeof(o));
    o.variant = variant;
    return o;
}

namespace {

// Host-pointer form shared by the three drop-in entry points.
// kth_only: knn_dist receives query_nb floats, the distance to each query's
// k-th neighbour (row k-1 of the full result), and knn_index is not used.
bool run_host(const float *ref, int R, const float *query, int Q, int dim, int k,
              float *knn_dist, int *knn_index, int variant, bool kth_only = false) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!check_args(ref, R, query, Q, dim, k, knn_dist, kth_only ? (const void *)knn_dist : knn_index))
        return false;
    if (!ensure_device()) return false;
    CallOrder order(nullptr);
    const knnhip_options o = default_options(variant);
    // Problems beyond the device memory budget stream through it.
    bool explicit_budget = false;
    const size_t budget = device_budget(&explicit_budget);
    if (search_footprint(R, Q, dim, k, o, false) > budget)
        return run_streamed(ref, R, query, Q, dim, k, knn_dist, knn_index, o, budget, kth_only) &&
               nonfinite_check(nullptr);
    Plan p = make_plan(R, Q, dim, k, o);
    // Large reference sets go up in chunks of whole slices, each scanned as
    // soon as it lands, so the copy overlaps the search.
    bool chunked = R >= kUploadMinRefs && p.use_scan && !p.lds && !p.bf3;
    if (chunked) {
        knnhip_options oc = o;
        oc.slices = std::max(p.nslices, kUploadChunks);
        p = make_plan(R, Q, dim, k, oc);
        chunked = p.use_scan && !p.lds && !p.bf3 && p.nslices >= 2;
    }
    if (chunked) {
        if (!g_ws.copy) {
            KNN_TRY(hipStreamCreateWithFlags(&g_ws.copy, hipStreamNonBlocking), "stream create");
            for (auto &e : g_ws.up_ev) KNN_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event create");
            KNN_TRY(hipEventCreateWithFlags(&g_ws.qry_ev, hipEventDisableTiming), "event create");
        }
    }
    g_last_plan = plan_json(p, R, Q, dim, k);
    g_last_plan.pop_back();
    g_last_plan += ", \"variant\": " + std::to_string(variant) + "}";
    if (chunked) {
        g_last_plan.pop_back();
        g_last_plan += ", \"upload_chunks\": " + std::to_string(std::min(kUploadChunks, p.nslices)) + "}";
    }
    float *ref_d = (float *)ws_get(kRefPad, (size_t)dim * R * sizeof(float));
    float *qry_d = (float *)ws_get(kQryPad, (size_t)dim * Q * sizeof(float));
    float *dist_d = (float *)ws_get(kOutDist, (size_t)k * Q * sizeof(float));
    int *idx_d = (int *)ws_get(kOutIdx, (size_t)k * Q * sizeof(int));
    if (!ref_d || !qry_d || !dist_d || !idx_d) {
        set_error("Memory allocation error");
        return false;
    }
    hipStream_t s = nullptr;
    if (chunked) {
        // queries first on the upload stream, then the reference chunks
        if (!host_upload(qry_d, (long long)dim * Q, query, (long long)dim * Q, (long long)dim * Q, 1,
                         g_ws.copy))
            return false;
        KNN_TRY(hipEventRecord(g_ws.qry_ev, g_ws.copy), "event");
        KNN_TRY(hipStreamWaitEvent(s, g_ws.qry_ev, 0), "event wait");
    } else {
        KNN_TRY(hipMemcpy(qry_d, query, (size_t)dim * Q * sizeof(float), hipMemcpyHostToDevice),
                "Unable to copy data from host to device");
    }
    if (chunked) {
        const Upload up{ref, R, g_ws.copy, kUploadChunks};
        if (!run_device(ref_d, R, R, qry_d, Q, Q, dim, k, dist_d, idx_d, p, o, s, 0, &up)) {
            // copies already queued still read the caller's buffer: let them finish
            (void)hipStreamSynchronize(g_ws.copy);
            return false;
        }
    } else {
        KNN_TRY(hipMemcpy(ref_d, ref, (size_t)dim * R * sizeof(float), hipMemcpyHostToDevice),
                "Unable to copy data from host to device");
        if (!run_device(ref_d, R, R, qry_d, Q, Q, dim, k, dist_d, idx_d, p, o, s)) return false;
    }
    // synchronous copies on the null stream, behind the search
    if (kth_only) {
        KNN_TRY(hipMemcpy(knn_dist, dist_d + (size_t)(k - 1) * Q, (size_t)Q * sizeof(float),
                          hipMemcpyDeviceToHost),
                "Unable to copy data from device to host");
        return nonfinite_check(nullptr);
    }
    KNN_TRY(hipMemcpy(knn_dist, dist_d, (size_t)k * Q * sizeof(float), hipMemcpyDeviceToHost),
            "Unable to copy data from device to host");
    KNN_TRY(hipMemcpy(knn_index, idx_d, (size_t)k * Q * sizeof(int), hipMemcpyDeviceToHost),
            "Unable to copy data from device to host");
    return nonfinite_check(nullptr);
}

}  // namespace

bool report(bool ok) {
    if (!ok) {
        printf("ERROR: %s\n", g_last_error.c_str());
        fflush(stdout);
    }
    return ok;
}

}  // namespace knnhip

using namespace knnhip;

bool knn_cuda_global(const float *ref, int ref_nb, const float *query, int query_nb, int dim,
                     int k, float *knn_dist, int *knn_index) {
    return report(run_host(ref, ref_nb, query, query_nb, dim, k, knn_dist, knn_index, 0));
}

bool knn_cuda_texture(const float *ref, int ref_nb, const float *query, int query_nb, int dim,
                      int k, float *knn_dist, int *knn_index) {
    return report(run_host(ref, ref_nb, query, query_nb, dim, k, knn_dist, knn_index, 1));
}

bool knn_cublas(const float *ref, int ref_nb, const float *query, int query_nb, int dim, int k,
                float *knn_dist, int *knn_index) {
    return report(run_host(ref, ref_nb, query, query_nb, dim, k, knn_dist, knn_index, 2));
}

extern "C" {

int knnhip_global(const float *ref, int ref_nb, const float *query, int query_nb, int dim, int k,
                  float *knn_dist, int *knn_index) {
    return knn_cuda_global(ref, ref_nb, query, query_nb, dim, k, knn_dist, knn_index) ? 1 : 0;
}

int knnhip_texture(const float *ref, int ref_nb, const float *query, int query_nb, int dim, int k,
                   float *knn_dist, int *knn_index) {
    return knn_cuda_texture(ref, ref_nb, query, query_nb, dim, k, knn_dist, knn_index) ? 1 : 0;
}

int knnhip_cublas(const float *ref, int ref_nb, const float *query, int query_nb, int dim, int k,
                  float *knn_dist, int *knn_index) {
    return knn_cublas(ref, ref_nb, query, query_nb, dim, k, knn_dist, knn_index) ? 1 : 0;
}

// The k-th-neighbour distance of every query, as the reference's legacy
// programs without indexes return it (legacy/KNN.CUDA.without.indexes/
// knn_cuda_without_indexes.cu:278-398): kth_dist[q] = row k-1 of
// knn_cuda_global's distances, bit for bit.
int knnhip_kth_distance(const float *ref, int ref_nb, const float *query, int query_nb, int dim,
                        int k, float *kth_dist) {
    return report(run_host(ref, ref_nb, query, query_nb, dim, k, kth_dist, nullptr, 0, true)) ? 1 : 0;
}

int knnhip_search_device(const float *ref, int ref_nb, long long ref_ld, const float *query,
                         int query_nb, long long query_ld, int dim, int k, float *knn_dist,
                         int *knn_index, const knnhip_options *opt, void *stream) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!check_args(ref, ref_nb, query, query_nb, dim, k, knn_dist, knn_index)) return 0;
    if (ref_ld < ref_nb || query_ld < query_nb) {
        set_error("leading dimensions must be >= point counts");
        return 0;
    }
    if (!ensure_device()) return 0;
    knnhip_options o = opt ? *opt : default_options(0);
    hipStream_t s = (hipStream_t)stream;
    CallOrder order(s);
    const Plan p = make_plan(ref_nb, query_nb, dim, k, o);
    g_last_plan = plan_json(p, ref_nb, query_nb, dim, k);
    g_last_plan.pop_back();
    g_last_plan += ", \"variant\": " + std::to_string(o.variant) + "}";
    return run_device(ref, ref_ld, ref_nb, query, query_ld, query_nb, dim, k, knn_dist, knn_index,
                      p, o, s)
               ? 1
               : 0;
}

}  // extern "C"

namespace {

// Scratch of knnhip_search_device_noalloc: every buffer its search needs
// besides the caller's inputs and outputs (norms, packed tiles, point-major
// rows, heaps, thresholds, counters, exact-scan lists).  Each device the
// library runs on holds one copy until the process ends.  In host code the
// array's name is only a stand-in: its device address comes from
// hipGetSymbolAddress, once per device, and only that address is used.
constexpr size_t kNoAllocBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices] = {};

// The current device's scratch address (nullptr with an error set).
void *noalloc_scratch() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        set_error("cannot query the current device");
        return nullptr;
    }
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            set_error("scratch address: %s", hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

int knnhip_search_device_noalloc(const float *ref, int ref_nb, long long ref_ld, const float *query,
                                 int query_nb, long long query_ld, int dim, int k, float *knn_dist,
                                 int *knn_index, const knnhip_options *opt, void *stream) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!check_args(ref, ref_nb, query, query_nb, dim, k, knn_dist, knn_index)) return 0;
    if (ref_ld < ref_nb || query_ld < query_nb) {
        set_error("leading dimensions must be >= point counts");
        return 0;
    }
    if (!ensure_device()) return 0;
    knnhip_options o = opt ? *opt : default_options(0);
    o.timing = 0;  // (its events would be created on first use)
    hipStream_t s = (hipStream_t)stream;
    const Plan p = make_plan(ref_nb, query_nb, dim, k, o);
    const size_t need = run_device_bytes(p, ref_nb, query_nb, dim, k);
    if (need > kNoAllocBytes) {
        set_error("knnhip_search_device_noalloc: this search needs %zu bytes of scratch, more than the "
                  "%zu it has (use knnhip_search_device)",
                  need, kNoAllocBytes);
        return 0;
    }
    void *base = noalloc_scratch();
    if (!base) return 0;
    CallOrder order(s);
    g_last_plan = plan_json(p, ref_nb, query_nb, dim, k);
    Arena arena;
    arena.base = static_cast<unsigned char *>(base);
    arena.cap = kNoAllocBytes;
    g_arena = &arena;
    const bool ok = run_device(ref, ref_ld, ref_nb, query, query_ld, query_nb, dim, k, knn_dist, knn_index,
                               p, o, s);
    g_arena = nullptr;
    return ok ? 1 : 0;
}

int knnhip_last_fallback_count(void) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_last_fb) return 0;
    int n = -1;
    unsigned bad = 0;
    if (hipStreamSynchronize(g_last_stream) != hipSuccess ||
        hipMemcpy(&n, g_last_fb, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
        (g_last_nonfinite &&
         hipMemcpy(&bad, g_last_nonfinite, sizeof(unsigned), hipMemcpyDeviceToHost) != hipSuccess)) {
        set_error("cannot read the fallback count");
        return -1;
    }
    if (bad) {
        set_error("non-finite input: a reference or query point has a NaN or infinite squared norm");
        return -1;
    }
    return n;
}

float knnhip_last_stage_ms(int stage) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (stage < 0 || stage > 3) return -1.f;
    return g_stage_ms[stage];
}

const char *knnhip_last_error(void) { return g_last_error.c_str(); }

const char *knnhip_last_plan(void) { return g_last_plan.c_str(); }

int knnhip_generate_uniform(float *out, int n, int dim, unsigned long long seed, float scale,
                            void *stream) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!out || n < 1 || dim < 1) {
        set_error("bad generator arguments");
        return 0;
    }
    if (!ensure_device()) return 0;
    hipError_t e = launch_gen_uniform(out, n, dim, seed, scale, (hipStream_t)stream);
    if (e != hipSuccess) {
        set_error("generator launch: %s", hipGetErrorString(e));
        return 0;
    }
    return 1;
}

int knnhip_generate_mixture(float *out, int n, int dim, int ncomp, float sigma,
                            unsigned long long centre_seed, unsigned long long point_seed,
                            void *stream) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!out || n < 1 || dim < 1 || ncomp < 1) {
        set_error("bad generator arguments");
        return 0;
    }
    if (!ensure_device()) return 0;
    hipError_t e = launch_gen_mixture(out, n, dim, ncomp, sigma, centre_seed, point_seed,
                                      (hipStream_t)stream);
    if (e != hipSuccess) {
        set_error("generator launch: %s", hipGetErrorString(e));
        return 0;
    }
    return 1;
}

const char *knnhip_version(void) { return "knnhip 0.2 (gfx950)"; }

}  // extern "C"
