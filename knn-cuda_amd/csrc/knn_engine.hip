// Host side of the exact k-NN engine: argument checks, launch planning,
// device workspace, and the drop-in entry points.
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
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/knncuda.h"
#include "knn_common.h"
#include "knn_index.h"
#include "knn_router.h"

namespace knnhip {

// Owned device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    bool alloc(size_t bytes) {
        release();
        if (bytes == 0) bytes = 256;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            return false;
        }
        n = bytes;
        return true;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
    template <class T> T *as() const { return reinterpret_cast<T *>(p); }
};

// Launchers defined in knn_scan.hip / knn_select.hip.
int scan_pick_ks(int dim);
size_t scan_lds_bytes(int kp, int waves);
int scan_blocks_per_cu(int ks, bool multi, int waves, int kp);
hipError_t launch_scan(const ScanArgs &a, int ks, bool multi, int waves, hipStream_t s);
int scan_bf3_nch(int dim);
size_t scan_bf3_lds_bytes(int kp);
hipError_t launch_scan_lds(const ScanArgs &a, int nch, bool bf3, hipStream_t s);
hipError_t launch_norms(const NormArgs &a, hipStream_t s);
hipError_t launch_select(const FinalArgs &a, hipStream_t s);
hipError_t launch_exact(const ExactArgs &a, int grid, hipStream_t s);
hipError_t launch_gen_uniform(float *out, int n, int dim, uint64_t seed, float scale,
                              hipStream_t s);
hipError_t launch_gen_mixture(float *out, int n, int dim, int ncomp, float sigma,
                              uint64_t centre_seed, uint64_t point_seed, hipStream_t s);

namespace {

// Candidates kept beyond k (kp = k + margin): the certificate needs the
// kp-th approximate key to clear the k-th exact distance by the FP32 error
// bound.  Large k goes with dense neighbourhoods, so it starts wider.
constexpr int kDefaultMargin = 16;
constexpr int kIndexMargin = 8;  // block-index searches: fewer heap insertions per block
constexpr int kDefaultMarginLargeK = 64;
constexpr int kLargeK = 64;
// Certificate failures are retried through the scan with wider margins
// (kEscalationMargins[level - 1]) before the exact per-query scan.
constexpr int kEscalationMargins[] = {64, 256};
constexpr int kMaxEscalation = 2;
int default_margin(int k) { return k >= kLargeK ? kDefaultMarginLargeK : kDefaultMargin; }
constexpr int kExactGrid = 512;
constexpr size_t kScanLdsMax = 160 * 1024;

std::mutex g_mu;
std::string g_last_error;
float g_stage_ms[4] = {-1.f, -1.f, -1.f, -1.f};
hipStream_t g_last_stream = nullptr;
const int *g_last_fb = nullptr;  // device counter of the last search's exact-scan queries
std::string g_last_plan;
long long g_batch_fail = 0;  // certificate failures of the last batched search
int g_cu_count = 0;

void set_error(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

// Grow-only device buffers, one per role.
enum Slot {
    kRefPad, kQryPad, kRNorm, kQNorm, kRMax, kPartVal, kPartIdx, kFbCount, kFbList,
    kFbKey, kFbIdx, kOutDist, kOutIdx, kRefPm, kQryPm, kPacked, kTheta, kPackedBf, kNumSlots
};
struct Workspace {
    void *ptr[kNumSlots] = {};
    size_t cap[kNumSlots] = {};
    hipEvent_t ev[5] = {};
    bool have_events = false;
};
Workspace g_ws;

void *ws_get(Slot s, size_t bytes) {
    if (bytes == 0) bytes = 256;
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

// Launch plan of one call.
struct Plan {
    bool use_scan = false;  // false: exact full scan for every query
    int ks = 0;
    bool multi = false;
    int nchunks = 1;
    int dpad = 0;
    int waves = 4;
    int kp = 0;
    int nslices = 1;
    int slice_len = 0;
    int nqb = 0;
    int xcd = 0;
    int maxc = 0;
    int blocks_per_cu = 0;
    int cus = 0;
    int r_pad = 0;            // references rounded up to whole scan steps
    int oct = 0;              // dimension octets of the packed layout
    long long slice_span = 0; // packed bytes per slice
    bool lds = false;         // 128-query scan with LDS-shared reference chunks
    bool bf3 = false;         // ... with bfloat16 three-term products
    int nch = 0;              // its 16-dimension chunks
};

// Queries per scan block: 32 for the per-wave-tile FP32 scan, 128 for the
// LDS-shared scan.
int block_queries(const Plan &p) { return p.lds ? 128 : kQueriesPerWave; }

// Cost model of one slice count: the scan runs in `rounds` waves of
// resident blocks, each block's four waves sharing ceil(steps / S) reference
// steps, plus a warm-up of about two steps while the heaps fill.
double slice_cost(int S, int steps, int nqb, int resident) {
    const long long units = (long long)nqb * S;
    const long long rounds = (units + resident - 1) / resident;
    return (double)rounds * ((double)((steps + S - 1) / S) / 4.0 + 2.0);
}

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
        p.multi = false;
    } else {
        p.ks = 64;
        p.multi = true;
        p.nchunks = (dim + 2 * p.ks - 1) / (2 * p.ks);
    }
    p.dpad = 2 * p.ks * p.nchunks;
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
    p.blocks_per_cu = std::max(1, scan_blocks_per_cu(p.ks, p.multi, p.waves, p.kp));
    const int resident = p.cus * p.blocks_per_cu;
    int s = smin;
    if (o.slices > 0) {
        s = std::max(smin, std::min(o.slices, smax));
    } else {
        double best = slice_cost(smin, steps, p.nqb, resident);
        for (int c = smin + 1; c <= smax; ++c) {
            const double cost = slice_cost(c, steps, p.nqb, resident);
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
    if (o.precision == 1 || o.block_queries == 128) {
        // LDS-shared scan: 128 queries per block, one block per CU; slices
        // fill whole rounds of CUs (the same cost model, resident = CUs).
        const int nch = scan_bf3_nch(dim);
        if (nch && scan_bf3_lds_bytes(p.kp) <= kScanLdsMax) {
            p.lds = true;
            p.bf3 = o.precision == 1;
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
                double best = slice_cost(bs, steps, p.nqb, p.cus) ;
                for (int c = bs + 1; c <= smax; ++c) {
                    const double cost = slice_cost(c, steps, p.nqb, p.cus);
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
             "\"ks\": %d, \"chunks\": %d, \"dpad\": %d, \"waves_per_block\": %d, \"kp\": %d, "
             "\"slices\": %d, \"slice_len\": %d, \"query_blocks\": %d, \"xcd_map\": %d, "
             "\"blocks_per_cu\": %d, \"cus\": %d, \"max_candidates\": %d, \"precision\": \"%s\", \"block_queries\": %d}",
             R, Q, dim, k, p.use_scan ? "true" : "false", p.ks, p.nchunks, p.dpad, p.waves, p.kp,
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

#define KNN_TRY(expr, what)                                                        \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess) {                                                    \
            set_error("%s: %s", what, hipGetErrorString(_e));                      \
            return false;                                                          \
        }                                                                          \
    } while (0)

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

// Core: device pointers, dim rows of ldr / ldq floats.
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

// Buffers of one escalation level: the failed queries' ids, their columns
// and their results.
struct Escalation {
    DevBuf list, qsub, dsub, isub;
};
Escalation g_esc[kMaxEscalation];

bool run_device(const float *ref, long long ldr, int R, const float *query, long long ldq, int Q,
                int dim, int k, float *out_dist, int *out_idx, const Plan &p,
                const knnhip_options &o, hipStream_t s, int level = 0) {
    const bool timing = o.timing != 0;
    if (timing && !g_ws.have_events) {
        for (auto &e : g_ws.ev) KNN_TRY(hipEventCreate(&e), "event create");
        g_ws.have_events = true;
    }
    for (float &m : g_stage_ms) m = -1.f;
    int *fb_count = (int *)ws_get(kFbCount, sizeof(int));
    if (!fb_count) {
        set_error("Memory allocation error");
        return false;
    }
    KNN_TRY(hipMemsetAsync(fb_count, 0, sizeof(int), s), "memset");

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
    const int egrid = std::min(kExactGrid, Q);
    ea.scratch_key = (float *)ws_get(kFbKey, (size_t)egrid * kWave * k * sizeof(float));
    ea.scratch_idx = (int *)ws_get(kFbIdx, (size_t)egrid * kWave * k * sizeof(int));
    if (!ea.scratch_key || !ea.scratch_idx) {
        set_error("Memory allocation error");
        return false;
    }

    if (!p.use_scan) {
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
        if (p.bf3) {
            na.packed_bf = (unsigned short *)packed;
            na.nch_bf = p.nch;
        } else
            na.packed = (float *)packed;
        KNN_TRY(launch_norms(na, s), "norm kernel launch");
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
        if (p.lds)
            KNN_TRY(launch_scan_lds(sa, p.nch, p.bf3, s), "scan kernel launch");
        else
            KNN_TRY(launch_scan(sa, p.ks, p.multi, p.waves, s), "scan kernel launch");
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
        KNN_TRY(launch_select(fa, s), "select kernel launch");
        if (timing) KNN_TRY(hipEventRecord(g_ws.ev[3], s), "event");

        // Queries that failed the certificate: once more through the scan
        // with a wider margin while one is left, else the exact per-query
        // scan (whose grid reads the count on the device).  A forced
        // fallback goes to the exact scan directly.
        const int cur_margin = p.kp - k;
        int next_margin = 0;
        for (int l = level; l < kMaxEscalation && !next_margin && !o.force_fallback; ++l)
            if (kEscalationMargins[l] > cur_margin && k + cur_margin < R) next_margin = kEscalationMargins[l];
        int nfail = 0;
        if (next_margin) {
            KNN_TRY(hipMemcpyAsync(&nfail, fb_count, sizeof(int), hipMemcpyDeviceToHost, s), "copy");
            KNN_TRY(hipStreamSynchronize(s), "sync");
        }
        if (next_margin && nfail > 0) {
            Escalation &e = g_esc[level];
            auto need = [](DevBuf &b, size_t bytes) { return b.n >= bytes || b.alloc(bytes); };
            if (!need(e.list, (size_t)nfail * 4) || !need(e.qsub, (size_t)nfail * dim * 4) ||
                !need(e.dsub, (size_t)nfail * k * 4) || !need(e.isub, (size_t)nfail * k * 4)) {
                set_error("Memory allocation error");
                return false;
            }
            KNN_TRY(hipMemcpyAsync(e.list.p, fb_list, (size_t)nfail * 4, hipMemcpyDeviceToDevice, s),
                    "copy");
            KNN_TRY(idx_gather_columns(query, ldq, dim, e.list.as<int>(), nfail, e.qsub.as<float>(), s),
                    "gather");
            knnhip_options o2 = o;
            o2.margin = next_margin;
            o2.force_fallback = 0;
            o2.timing = 0;
            o2.precision = 0;  // the rescans use FP32 products
            const Plan p2 = make_plan(R, nfail, dim, k, o2);
            if (!run_device(ref, ldr, R, e.qsub.as<float>(), nfail, nfail, dim, k, e.dsub.as<float>(),
                            e.isub.as<int>(), p2, o2, s, level + 1))
                return false;
            KNN_TRY(idx_scatter_results(e.dsub.as<float>(), e.isub.as<int>(), k, nfail,
                                        e.list.as<int>(), Q, out_dist, out_idx, s),
                    "scatter");
        } else if (!next_margin) {
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

knnhip_options default_options(int variant) {
    knnhip_options o;
    memset(&o, 0, sizeof(o));
    o.variant = variant;
    return o;
}

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
    const knnhip_options o = default_options(variant);
    const Plan p = make_plan(R, Q, dim, k, o);
    g_last_plan = plan_json(p, R, Q, dim, k);
    float *ref_d = (float *)ws_get(kRefPad, (size_t)dim * R * sizeof(float));
    float *qry_d = (float *)ws_get(kQryPad, (size_t)dim * Q * sizeof(float));
    float *dist_d = (float *)ws_get(kOutDist, (size_t)k * Q * sizeof(float));
    int *idx_d = (int *)ws_get(kOutIdx, (size_t)k * Q * sizeof(int));
    if (!ref_d || !qry_d || !dist_d || !idx_d) {
        set_error("Memory allocation error");
        return false;
    }
    hipStream_t s = nullptr;
    KNN_TRY(hipMemcpyAsync(ref_d, ref, (size_t)dim * R * sizeof(float), hipMemcpyHostToDevice, s),
            "Unable to copy data from host to device");
    KNN_TRY(hipMemcpyAsync(qry_d, query, (size_t)dim * Q * sizeof(float), hipMemcpyHostToDevice,
                           s),
            "Unable to copy data from host to device");
    if (!run_device(ref_d, R, R, qry_d, Q, Q, dim, k, dist_d, idx_d, p, o, s)) return false;
    if (kth_only) {
        KNN_TRY(hipMemcpyAsync(knn_dist, dist_d + (size_t)(k - 1) * Q, (size_t)Q * sizeof(float),
                               hipMemcpyDeviceToHost, s),
                "Unable to copy data from device to host");
        KNN_TRY(hipStreamSynchronize(s), "Unable to execute kernel");
        return true;
    }
    KNN_TRY(hipMemcpyAsync(knn_dist, dist_d, (size_t)k * Q * sizeof(float), hipMemcpyDeviceToHost,
                           s),
            "Unable to copy data from device to host");
    KNN_TRY(hipMemcpyAsync(knn_index, idx_d, (size_t)k * Q * sizeof(int), hipMemcpyDeviceToHost,
                           s),
            "Unable to copy data from device to host");
    KNN_TRY(hipStreamSynchronize(s), "Unable to execute kernel");
    return true;
}

bool report(bool ok) {
    if (!ok) {
        printf("ERROR: %s\n", g_last_error.c_str());
        fflush(stdout);
    }
    return ok;
}


// ---------------------------------------------------------------------------
// Block index (router form)
// ---------------------------------------------------------------------------

}  // namespace


}  // namespace knnhip

// The opaque index handle of the C API.
struct knnhip_index {
    int R = 0, dim = 0, nb = 0, pm_ld = 0, oct = 0, dpad = 0, ks = 0, nchunks = 1;
    bool multi = false;
    int n_pos = 0, r_pad = 0;
    const float *ref = nullptr;  // caller's dim-major references (exact fallback)
    long long ldr = 0;
    float build_ms = 0.f;
    knnhip::DevBuf packed, pm, norms, perm, block_start, block_count, rmax;
    // fine balls in block-sorted order: centre, squared norm of the centre,
    // radius (< 0: empty ball), block; block b's balls are fb_start[b] ..
    // fb_start[b+1]
    int nf = 0;
    knnhip::DevBuf fcent, fnorm, frad, fblock, fb_start;
    // last search statistics
    int last_probes = 0;
    int last_fail = 0;
    long long last_tasks = 0;
    float last_ms[6] = {-1, -1, -1, -1, -1, -1};
    float rank_ms[3] = {-1, -1, -1};  // query norms, block bounds, block choice
};

namespace knnhip {
namespace {

// Lloyd iterations on the point-major rows `pm` (n points) from the `k`
// centroids already in `cent` (point-major, ld pm_ld).  The assignment step
// is the engine itself (centroids as the reference set, k = 1); the final
// assignment is left in `assign`, its histogram in `counts`.
bool kmeans(const float *pm, const float *dm, long long ldd, int n, int dim, int pm_ld, int k,
            int iters, float *cent, int *assign, int *counts, hipStream_t s) {
    DevBuf cent_dm, adist, start, cursor, order;
    if (!cent_dm.alloc((size_t)k * dim * 4) || !adist.alloc((size_t)n * 4) ||
        !start.alloc((size_t)(k + 1) * 4) || !cursor.alloc((size_t)(k + 1) * 4) ||
        !order.alloc((size_t)n * 4)) {
        set_error("Memory allocation error");
        return false;
    }
    knnhip_options ao;
    memset(&ao, 0, sizeof(ao));
    ao.margin = 8;
    const Plan ap = make_plan(k, n, dim, 1, ao);
    for (int it = 0; it <= iters; ++it) {
        KNN_TRY(idx_transpose(cent, pm_ld, k, dim, cent_dm.as<float>(), s), "transpose");
        if (!run_device(cent_dm.as<float>(), k, k, dm, ldd, n, dim, 1, adist.as<float>(), assign,
                        ap, ao, s))
            return false;
        KNN_TRY(hipMemsetAsync(counts, 0, (size_t)k * 4, s), "memset");
        KNN_TRY(idx_histogram(assign, n, counts, s), "histogram");
        if (it == iters) break;
        KNN_TRY(idx_prefix(counts, k, 1, start.as<int>(), nullptr, s), "prefix");
        KNN_TRY(hipMemcpyAsync(cursor.p, start.p, (size_t)k * 4, hipMemcpyDeviceToDevice, s), "copy");
        KNN_TRY(idx_scatter(assign, n, cursor.as<int>(), order.as<int>(), s), "scatter");
        KNN_TRY(idx_centroids(pm, pm_ld, dim, order.as<int>(), start.as<int>(), counts, k, cent,
                              pm_ld, s),
                "centroids");
    }
    return true;
}

// Block index build.
//  1. Fine balls: k-means over the references with kBallsPerBlock * nb
//     centres (k-means++ seeding on an evenly strided sample of kSeedPerCentre
//     points per centre, then `iters` Lloyd steps with the engine as the
//     assignment step).  Each ball's radius bounds its members.
//  2. Blocks: nb seeds drawn by k-means++ over the ball centres; each ball
//     joins the nearest of its kGroupCandidates nearest seeds that stays
//     within kBlockSlack times the mean block size.
//  3. Layout: references block by block (each block starting at a multiple
//     of 32), each ball contiguous inside its block; balls sorted by block.
// A block's lower bound for a query is the smallest over its balls.
constexpr int kSeedPerCentre = 16;
constexpr int kGroupCandidates = 32;
constexpr double kBlockSlack = 1.25;
constexpr unsigned long long kBallSeed = 0x6B4E4E5EEDBA11Full;
constexpr unsigned long long kBlockSeed = 0x6B4E4E5EEDB10C5ull;

bool index_build(knnhip_index *ix, const float *ref, int R, long long ldr, int dim, int nb,
                 int iters, hipStream_t s) {
    hipEvent_t e0, e1;
    KNN_TRY(hipEventCreate(&e0), "event");
    KNN_TRY(hipEventCreate(&e1), "event");
    KNN_TRY(hipEventRecord(e0, s), "event");
    ix->R = R;
    ix->dim = dim;
    ix->nb = nb;
    ix->ref = ref;
    ix->ldr = ldr;
    ix->pm_ld = (dim + 3) & ~3;
    if (ix->pm_ld > kIndexMaxLd) {
        set_error("dimension too large for the block index (dim=%d, at most %d)", dim, kIndexMaxLd);
        return false;
    }
    knnhip_options o;
    memset(&o, 0, sizeof(o));
    {
        Plan p = make_plan(R, 32, dim, 1, o);
        if (!p.use_scan) {
            set_error("dimension too large for the block index");
            return false;
        }
        ix->ks = p.ks;
        ix->multi = p.multi;
        ix->nchunks = p.nchunks;
        ix->dpad = p.dpad;
        ix->oct = p.oct;
    }
    const int pm_ld = ix->pm_ld;
    const int nf = (int)std::min<long long>(R, (long long)kBallsPerBlock * nb);
    ix->nf = nf;
    const int m = (int)std::min<long long>(R, (long long)kSeedPerCentre * nf);
    const int kc = std::min(kGroupCandidates, nb);
    const int cap = (int)std::ceil(kBlockSlack * (double)R / (double)nb);
    const int mmax = std::max(m, nf);
    DevBuf pm0, norms0, fassign, fcount, fcent, bcent, fcent_dm, bcent_dm, cand_d, cand_i, fblock,
        bfine, cursor, fstart, fsorted, mind, cidx, wsum, ticket;
    if (!pm0.alloc((size_t)R * pm_ld * 4) || !norms0.alloc((size_t)R * 4) || !ix->rmax.alloc(4) ||
        !fassign.alloc((size_t)R * 4) || !fcount.alloc((size_t)nf * 4) ||
        !fcent.alloc((size_t)nf * pm_ld * 4) || !bcent.alloc((size_t)nb * pm_ld * 4) ||
        !fcent_dm.alloc((size_t)nf * dim * 4) || !bcent_dm.alloc((size_t)nb * dim * 4) ||
        !cand_d.alloc((size_t)kc * nf * 4) || !cand_i.alloc((size_t)kc * nf * 4) ||
        !fblock.alloc((size_t)nf * 4) || !bfine.alloc((size_t)nb * 4) ||
        !cursor.alloc((size_t)(nb + 1) * 4) || !fstart.alloc((size_t)nf * 4) ||
        !fsorted.alloc((size_t)nf * 4) || !mind.alloc((size_t)mmax * 4) ||
        !cidx.alloc((size_t)std::max(nf, nb) * 4) || !wsum.alloc((size_t)kKppGroups * 8) ||
        !ticket.alloc(4) || !ix->block_start.alloc((size_t)(nb + 1) * 4) ||
        !ix->block_count.alloc((size_t)nb * 4) || !ix->fcent.alloc((size_t)nf * pm_ld * 4) ||
        !ix->frad.alloc((size_t)nf * 8) || !ix->fnorm.alloc((size_t)nf * 4) ||
        !ix->fblock.alloc((size_t)nf * 4) || !ix->fb_start.alloc((size_t)(nb + 1) * 4)) {
        set_error("Memory allocation error");
        return false;
    }
    // point-major copy in the caller's order, norms, max norm
    KNN_TRY(hipMemsetAsync(ix->rmax.p, 0, 4, s), "memset");
    NormArgs na{};
    na.p = ref;
    na.ld = ldr;
    na.n = R;
    na.n_pad = R;
    na.dim = dim;
    na.oct = (dim + 7) / 8;
    na.out = norms0.as<float>();
    na.maxbits = ix->rmax.as<unsigned>();
    na.pm = pm0.as<float>();
    na.pm_ld = pm_ld;
    KNN_TRY(launch_norms(na, s), "norm kernel launch");
    // 1. fine balls
    KNN_TRY(idx_kpp_seed(pm0.as<float>(), pm_ld, dim, R, m, nf, kBallSeed, mind.as<float>(),
                         cidx.as<int>(), wsum.as<double>(), ticket.as<unsigned>(),
                         fcent.as<float>(), s),
            "k-means++ seeding");
    if (!kmeans(pm0.as<float>(), ref, ldr, R, dim, pm_ld, nf, iters, fcent.as<float>(),
                fassign.as<int>(), fcount.as<int>(), s))
        return false;
    // 2. blocks: seeds over the ball centres, then balanced grouping
    KNN_TRY(idx_kpp_seed(fcent.as<float>(), pm_ld, dim, nf, nf, nb, kBlockSeed, mind.as<float>(),
                         cidx.as<int>(), wsum.as<double>(), ticket.as<unsigned>(),
                         bcent.as<float>(), s),
            "k-means++ seeding");
    KNN_TRY(idx_transpose(fcent.as<float>(), pm_ld, nf, dim, fcent_dm.as<float>(), s), "transpose");
    KNN_TRY(idx_transpose(bcent.as<float>(), pm_ld, nb, dim, bcent_dm.as<float>(), s), "transpose");
    {
        knnhip_options ao;
        memset(&ao, 0, sizeof(ao));
        const Plan ap = make_plan(nb, nf, dim, kc, ao);
        if (!run_device(bcent_dm.as<float>(), nb, nb, fcent_dm.as<float>(), nf, nf, dim, kc,
                        cand_d.as<float>(), cand_i.as<int>(), ap, ao, s))
            return false;
    }
    KNN_TRY(idx_group_fine(cand_i.as<int>(), kc, nf, fcount.as<int>(), nb, cap, fblock.as<int>(), s),
            "grouping");
    // 3. layout
    KNN_TRY(hipMemsetAsync(ix->block_count.p, 0, (size_t)nb * 4, s), "memset");
    KNN_TRY(hipMemsetAsync(bfine.p, 0, (size_t)nb * 4, s), "memset");
    KNN_TRY(idx_block_count(fblock.as<int>(), fcount.as<int>(), nf, ix->block_count.as<int>(),
                            bfine.as<int>(), s),
            "block count");
    KNN_TRY(idx_prefix(ix->block_count.as<int>(), nb, kRefsPerTile, ix->block_start.as<int>(),
                       nullptr, s),
            "prefix");
    KNN_TRY(idx_prefix(bfine.as<int>(), nb, 1, ix->fb_start.as<int>(), nullptr, s), "prefix");
    int n_pos = 0;
    KNN_TRY(hipMemcpyAsync(&n_pos, ix->block_start.as<int>() + nb, 4, hipMemcpyDeviceToHost, s),
            "copy");
    KNN_TRY(hipStreamSynchronize(s), "sync");
    ix->n_pos = n_pos;
    ix->r_pad = (n_pos + kRefsPerStep - 1) / kRefsPerStep * kRefsPerStep;
    DevBuf slot_cursor;
    if (!ix->perm.alloc((size_t)ix->r_pad * 4) ||
        !ix->packed.alloc((size_t)ix->r_pad * ix->dpad * 4) ||
        !ix->pm.alloc((size_t)ix->r_pad * pm_ld * 4) || !ix->norms.alloc((size_t)ix->r_pad * 4) ||
        !slot_cursor.alloc((size_t)(nb + 1) * 4)) {
        set_error("Memory allocation error");
        return false;
    }
    KNN_TRY(hipMemcpyAsync(cursor.p, ix->block_start.p, (size_t)nb * 4, hipMemcpyDeviceToDevice, s),
            "copy");
    KNN_TRY(hipMemcpyAsync(slot_cursor.p, ix->fb_start.p, (size_t)nb * 4, hipMemcpyDeviceToDevice, s),
            "copy");
    KNN_TRY(idx_fine_start(fblock.as<int>(), fcount.as<int>(), nf, cursor.as<int>(),
                           fstart.as<int>(), slot_cursor.as<int>(), fsorted.as<int>(), s),
            "ball ranges");
    // positions: ball f's members go to fstart[f] .. fstart[f] + fcount[f]
    KNN_TRY(hipMemcpyAsync(fcent_dm.p, fstart.p, (size_t)nf * 4, hipMemcpyDeviceToDevice, s),
            "copy");  // fcent_dm is free now: scatter cursor
    KNN_TRY(hipMemsetD32Async((hipDeviceptr_t)ix->perm.p, -1, ix->r_pad, s), "memset");
    KNN_TRY(hipMemsetAsync(ix->pm.p, 0, (size_t)ix->r_pad * pm_ld * 4, s), "memset");
    KNN_TRY(idx_scatter(fassign.as<int>(), R, fcent_dm.as<int>(), ix->perm.as<int>(), s), "scatter");
    KNN_TRY(idx_gather_pack(pm0.as<float>(), pm_ld, dim, ix->oct, ix->perm.as<int>(), n_pos,
                            ix->r_pad, ix->packed.as<float>(), ix->pm.as<float>(),
                            ix->norms.as<float>(), s),
            "gather");
    // balls in block order: centres, radii over their member positions, norms
    KNN_TRY(idx_finalize_fine(ix->pm.as<float>(), pm_ld, dim, fcent.as<float>(), fsorted.as<int>(),
                              fstart.as<int>(), fcount.as<int>(), fblock.as<int>(), nf,
                              ix->fcent.as<float>(), ix->frad.as<double>(), ix->fblock.as<int>(), s),
            "balls");
    KNN_TRY(idx_row_norms(ix->fcent.as<float>(), pm_ld, nf, dim, ix->fnorm.as<float>(), s), "norms");
    KNN_TRY(hipEventRecord(e1, s), "event");
    KNN_TRY(hipEventSynchronize(e1), "event");
    KNN_TRY(hipEventElapsedTime(&ix->build_ms, e0, e1), "event");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return true;
}

// On-disk format of a block index (little-endian): magic "KNNHIPIX", uint32
// version (1), int32 R, dim, nb, nf, n_pos, r_pad; then block_start [nb+1],
// block_count [nb], perm [r_pad] (int32), ball centres [nf][pm_ld], centre
// norms [nf] (float32), radii [nf] (float64), ball blocks [nf] and ball
// starts [nb+1] (int32).  The packed tiles, point-major copy and norms are
// rebuilt from the caller's reference set when the file is loaded.
constexpr char kIndexMagic[8] = {'K', 'N', 'N', 'H', 'I', 'P', 'I', 'X'};
constexpr unsigned kIndexVersion = 1;

struct FileOut {
    FILE *f;
    bool ok = true;
    void put(const void *p, size_t n) { ok = ok && fwrite(p, 1, n, f) == n; }
    void put_dev(const DevBuf &b, size_t n) {
        std::vector<char> h(n);
        ok = ok && hipMemcpy(h.data(), b.p, n, hipMemcpyDeviceToHost) == hipSuccess;
        put(h.data(), n);
    }
};
struct FileIn {
    FILE *f;
    bool ok = true;
    void get(void *p, size_t n) { ok = ok && fread(p, 1, n, f) == n; }
    void get_dev(DevBuf &b, size_t n) {
        std::vector<char> h(n);
        get(h.data(), n);
        ok = ok && b.alloc(n) && hipMemcpy(b.p, h.data(), n, hipMemcpyHostToDevice) == hipSuccess;
    }
};

bool index_save(const knnhip_index *ix, const char *path) {
    FILE *f = fopen(path, "wb");
    if (!f) {
        set_error("cannot open %s for writing", path);
        return false;
    }
    FileOut o{f};
    const int hdr[6] = {ix->R, ix->dim, ix->nb, ix->nf, ix->n_pos, ix->r_pad};
    o.put(kIndexMagic, 8);
    o.put(&kIndexVersion, 4);
    o.put(hdr, sizeof(hdr));
    o.put_dev(ix->block_start, (size_t)(ix->nb + 1) * 4);
    o.put_dev(ix->block_count, (size_t)ix->nb * 4);
    o.put_dev(ix->perm, (size_t)ix->r_pad * 4);
    o.put_dev(ix->fcent, (size_t)ix->nf * ix->pm_ld * 4);
    o.put_dev(ix->fnorm, (size_t)ix->nf * 4);
    o.put_dev(ix->frad, (size_t)ix->nf * 8);
    o.put_dev(ix->fblock, (size_t)ix->nf * 4);
    o.put_dev(ix->fb_start, (size_t)(ix->nb + 1) * 4);
    const bool ok = o.ok && fclose(f) == 0;
    if (!ok) set_error("cannot write %s", path);
    return ok;
}

bool index_load(knnhip_index *ix, const char *path, const float *ref, int R, long long ldr, int dim,
                hipStream_t s) {
    FILE *f = fopen(path, "rb");
    if (!f) {
        set_error("cannot open %s", path);
        return false;
    }
    FileIn in{f};
    char magic[8] = {0};
    unsigned version = 0;
    int hdr[6] = {0};
    in.get(magic, 8);
    in.get(&version, 4);
    in.get(hdr, sizeof(hdr));
    if (!in.ok || memcmp(magic, kIndexMagic, 8) != 0 || version != kIndexVersion) {
        fclose(f);
        set_error("%s is not a block index file of version %u", path, kIndexVersion);
        return false;
    }
    if (hdr[0] != R || hdr[1] != dim || hdr[2] < 1 || hdr[3] < 1 || hdr[4] < R || hdr[5] < hdr[4]) {
        fclose(f);
        set_error("%s indexes %d points of dim %d, not %d of dim %d", path, hdr[0], hdr[1], R, dim);
        return false;
    }
    ix->R = R;
    ix->dim = dim;
    ix->nb = hdr[2];
    ix->nf = hdr[3];
    ix->n_pos = hdr[4];
    ix->r_pad = hdr[5];
    ix->ref = ref;
    ix->ldr = ldr;
    ix->pm_ld = (dim + 3) & ~3;
    knnhip_options o;
    memset(&o, 0, sizeof(o));
    const Plan p = make_plan(R, 32, dim, 1, o);
    if (!p.use_scan || ix->pm_ld > kIndexMaxLd) {
        fclose(f);
        set_error("dimension too large for the block index");
        return false;
    }
    ix->ks = p.ks;
    ix->multi = p.multi;
    ix->nchunks = p.nchunks;
    ix->dpad = p.dpad;
    ix->oct = p.oct;
    in.get_dev(ix->block_start, (size_t)(ix->nb + 1) * 4);
    in.get_dev(ix->block_count, (size_t)ix->nb * 4);
    in.get_dev(ix->perm, (size_t)ix->r_pad * 4);
    in.get_dev(ix->fcent, (size_t)ix->nf * ix->pm_ld * 4);
    in.get_dev(ix->fnorm, (size_t)ix->nf * 4);
    in.get_dev(ix->frad, (size_t)ix->nf * 8);
    in.get_dev(ix->fblock, (size_t)ix->nf * 4);
    in.get_dev(ix->fb_start, (size_t)(ix->nb + 1) * 4);
    char extra;
    const bool tail = fread(&extra, 1, 1, f) == 0;
    fclose(f);
    if (!in.ok || !tail) {
        set_error("%s is truncated or has trailing data", path);
        return false;
    }
    // rebuild the layouts from the references
    DevBuf pm0, norms0;
    if (!pm0.alloc((size_t)R * ix->pm_ld * 4) || !norms0.alloc((size_t)R * 4) || !ix->rmax.alloc(4) ||
        !ix->packed.alloc((size_t)ix->r_pad * ix->dpad * 4) ||
        !ix->pm.alloc((size_t)ix->r_pad * ix->pm_ld * 4) || !ix->norms.alloc((size_t)ix->r_pad * 4)) {
        set_error("Memory allocation error");
        return false;
    }
    KNN_TRY(hipMemsetAsync(ix->rmax.p, 0, 4, s), "memset");
    NormArgs na{};
    na.p = ref;
    na.ld = ldr;
    na.n = R;
    na.n_pad = R;
    na.dim = dim;
    na.oct = (dim + 7) / 8;
    na.out = norms0.as<float>();
    na.maxbits = ix->rmax.as<unsigned>();
    na.pm = pm0.as<float>();
    na.pm_ld = ix->pm_ld;
    KNN_TRY(launch_norms(na, s), "norm kernel launch");
    KNN_TRY(hipMemsetAsync(ix->pm.p, 0, (size_t)ix->r_pad * ix->pm_ld * 4, s), "memset");
    KNN_TRY(idx_gather_pack(pm0.as<float>(), ix->pm_ld, dim, ix->oct, ix->perm.as<int>(), ix->n_pos,
                            ix->r_pad, ix->packed.as<float>(), ix->pm.as<float>(), ix->norms.as<float>(),
                            s),
            "gather");
    KNN_TRY(hipStreamSynchronize(s), "sync");
    return true;
}

// Workspace of index searches (grown on demand, kept between calls).
struct IndexWork {
    DevBuf qnorm, qpm, lb, chosen, lb_rest, pcount, qstart, cursor, qlist, rows, tasks,
        pval, pidx, theta, fb_count, fb_list, qsub, dsub, isub;
    bool need(DevBuf &b, size_t bytes) { return b.n >= bytes || b.alloc(bytes); }
};
IndexWork g_iw, g_iw_retry;

// Probes of the retry that queries failing their certificate get before the
// full scan (ranked by the lower bounds alone).
constexpr int kRetryProbes = 32;

// depth 0: a caller's search; depth 1: the retry of its failed queries.
bool index_search(knnhip_index *ix, IndexWork &w, int depth, const float *query, int Q,
                  long long ldq, int k, int P, const float *router_scores, float *out_dist,
                  int *out_idx, const knnhip_options &o, hipStream_t s) {
    const int nb = ix->nb, dim = ix->dim, pm_ld = ix->pm_ld;
    const int margin = o.margin > 0 ? o.margin : (k >= kLargeK ? kDefaultMarginLargeK : kIndexMargin);
    const int kp = std::max(k, std::min(k + margin, ix->R));
    if (scan_lds_bytes(kp, 4) > kScanLdsMax || (long long)P * kp > kMaxUnion) {
        set_error("k or probe count too large for the block index (k=%d probes=%d)", k, P);
        return false;
    }
    P = std::max(1, std::min(P, nb));
    const long long npairs = (long long)Q * P;
    // Two scan phases: each query's first-ranked block, then its other
    // probes, which start from the thresholds the first phase left.
    const int nphase = P > 1 ? 2 : 1;
    const int pr[3] = {0, 1, P};
    int max_t[2] = {0, 0};
    for (int ph = 0; ph < nphase; ++ph)
        max_t[ph] = (int)(((long long)Q * (pr[ph + 1] - pr[ph]) + kQueriesPerWave - 1) /
                          kQueriesPerWave) + nb;
    const int max_tasks = max_t[0] + max_t[1];
    if (!w.need(w.qnorm, (size_t)Q * 4) || !w.need(w.qpm, (size_t)Q * pm_ld * 4) ||
        !w.need(w.lb, (size_t)Q * nb * 4) ||
        !w.need(w.chosen, (size_t)npairs * 4) || !w.need(w.lb_rest, (size_t)Q * 4) ||
        !w.need(w.pcount, (size_t)nb * 4) || !w.need(w.qstart, (size_t)(nb + 1) * 4) ||
        !w.need(w.cursor, (size_t)(nb + 1) * 4) ||
        !w.need(w.qlist, (size_t)max_tasks * kQueriesPerWave * 4) ||
        !w.need(w.rows, (size_t)npairs * 4) || !w.need(w.tasks, (size_t)max_tasks * sizeof(ScanTask)) ||
        !w.need(w.pval, (size_t)max_tasks * kQueriesPerWave * kp * 4) ||
        !w.need(w.pidx, (size_t)max_tasks * kQueriesPerWave * kp * 4) ||
        !w.need(w.theta, (size_t)Q * 4) || !w.need(w.fb_count, 4) || !w.need(w.fb_list, (size_t)Q * 4)) {
        set_error("Memory allocation error");
        return false;
    }
    hipEvent_t ev[6], evr[2];
    const bool timing = o.timing != 0;
    if (timing)
        for (auto &e : evr) KNN_TRY(hipEventCreate(&e), "event");
    if (timing)
        for (auto &e : ev) KNN_TRY(hipEventCreate(&e), "event");
    if (timing) KNN_TRY(hipEventRecord(ev[0], s), "event");
    // queries: norms + point-major copy
    NormArgs nq{};
    nq.p = query;
    nq.ld = ldq;
    nq.n = Q;
    nq.n_pad = Q;
    nq.dim = dim;
    nq.oct = (dim + 7) / 8;
    nq.out = w.qnorm.as<float>();
    nq.pm = w.qpm.as<float>();
    nq.pm_ld = pm_ld;
    KNN_TRY(launch_norms(nq, s), "norm kernel launch");
    if (timing) KNN_TRY(hipEventRecord(evr[0], s), "event");
    // block lower bounds (smallest over each block's balls), block-major;
    // without a router the ranking is by the negated bound (nearest first)
    KNN_TRY(idx_ball_lb(w.qpm.as<float>(), pm_ld, Q, dim, w.qnorm.as<float>(), ix->fcent.as<float>(),
                        ix->nf, ix->fnorm.as<float>(), ix->frad.as<double>(), ix->fblock.as<int>(),
                        ix->fb_start.as<int>(), nb, w.lb.as<float>(), s),
            "block bounds");
    const float *score = router_scores;
    if (timing) KNN_TRY(hipEventRecord(evr[1], s), "event");
    KNN_TRY(idx_pick_blocks(score, Q, nb, P, w.lb.as<float>(), ix->block_count.as<int>(),
                            w.chosen.as<int>(), w.lb_rest.as<float>(), s),
            "pick blocks");
    if (timing) KNN_TRY(hipEventRecord(ev[1], s), "event");
    if (timing) KNN_TRY(hipEventRecord(ev[2], s), "event");
    KNN_TRY(hipMemsetD32Async((hipDeviceptr_t)w.theta.p, 0xFF800000u, Q, s), "memset");
    KNN_TRY(hipMemsetAsync(w.fb_count.p, 0, 4, s), "memset");
    int tbase = 0;
    for (int ph = 0; ph < nphase; ++ph) {
        // (query, block) pairs of this phase -> per-block query lists -> tasks
        const int base = tbase * kQueriesPerWave;
        KNN_TRY(hipMemsetAsync(w.pcount.p, 0, (size_t)nb * 4, s), "memset");
        KNN_TRY(idx_pair_count(w.chosen.as<int>(), Q, P, pr[ph], pr[ph + 1], w.pcount.as<int>(), s),
                "pairs");
        KNN_TRY(idx_prefix(w.pcount.as<int>(), nb, kQueriesPerWave, w.qstart.as<int>(), nullptr, s),
                "prefix");
        KNN_TRY(hipMemcpyAsync(w.cursor.p, w.qstart.p, (size_t)nb * 4, hipMemcpyDeviceToDevice, s),
                "copy");
        KNN_TRY(idx_pair_scatter(w.chosen.as<int>(), Q, P, pr[ph], pr[ph + 1], w.cursor.as<int>(),
                                 w.qlist.as<int>() + base, w.rows.as<int>(), base, s),
                "pairs");
        KNN_TRY(idx_make_tasks(w.qstart.as<int>(), w.pcount.as<int>(), nb, ix->block_start.as<int>(),
                               ix->block_count.as<int>(), w.tasks.as<ScanTask>() + tbase, max_t[ph],
                               s),
                "tasks");
        ScanArgs sa{};
        sa.packed = ix->packed.as<float>();
        sa.oct = ix->oct;
        sa.R = ix->n_pos;
        sa.dim = dim;
        sa.query = query;
        sa.ldq = ldq;
        sa.Q = Q;
        sa.rnorm = ix->norms.as<float>();
        sa.theta = w.theta.as<unsigned>();
        sa.slice_len = ix->r_pad;
        sa.slice_span = (long long)ix->r_pad * ix->dpad * 4;
        sa.r_pad = ix->r_pad;
        sa.nslices = 1;
        sa.kp = kp;
        sa.nqb = 1;
        sa.nchunks = ix->nchunks;
        sa.xcd_map = 0;
        sa.part_val = w.pval.as<float>() + (size_t)base * kp;
        sa.part_idx = w.pidx.as<int>() + (size_t)base * kp;
        sa.tasks = w.tasks.as<ScanTask>() + tbase;
        sa.ntasks = max_t[ph];
        sa.qlist = w.qlist.as<int>() + base;
        sa.query_pm = w.qpm.as<float>();
        sa.query_pm_ld = pm_ld;
        KNN_TRY(launch_scan(sa, ix->ks, ix->multi, 4, s), "scan kernel launch");
        tbase += max_t[ph];
    }
    if (timing) KNN_TRY(hipEventRecord(ev[3], s), "event");
    // exact re-rank + certificate (skipped blocks bounded by lb_rest)
    FinalArgs fa{};
    fa.ref = ix->ref;
    fa.ldr = ix->ldr;
    fa.R = ix->R;
    fa.query = query;
    fa.ldq = ldq;
    fa.Q = Q;
    fa.dim = dim;
    fa.k = k;
    fa.kp = kp;
    fa.nrows = P;
    fa.rows = w.rows.as<int>();
    fa.perm = ix->perm.as<int>();
    fa.maxc = std::min(kMaxCand, std::max(2 * kp, 256));
    fa.part_val = w.pval.as<float>();
    fa.part_idx = w.pidx.as<int>();
    fa.ref_pm = ix->pm.as<float>();
    fa.query_pm = w.qpm.as<float>();
    fa.pm_ld = pm_ld;
    fa.qnorm = w.qnorm.as<float>();
    fa.rmax = ix->rmax.as<unsigned>();
    fa.theta = w.theta.as<unsigned>();
    fa.lb_rest = w.lb_rest.as<float>();
    fa.force_fallback = o.force_fallback;
    fa.out_dist = out_dist;
    fa.out_idx = out_idx;
    fa.fb_count = w.fb_count.as<int>();
    fa.fb_list = w.fb_list.as<int>();
    KNN_TRY(launch_select(fa, s), "select kernel launch");
    if (timing) KNN_TRY(hipEventRecord(ev[4], s), "event");
    // queries whose certificate failed: once more through the index with
    // more probes, then (or at once, on a retry) the plain full scan
    int nfail = 0;
    KNN_TRY(hipMemcpyAsync(&nfail, w.fb_count.p, 4, hipMemcpyDeviceToHost, s), "copy");
    KNN_TRY(hipStreamSynchronize(s), "sync");
    if (nfail > 0) {
        DevBuf &qsub = w.qsub, &dsub = w.dsub, &isub = w.isub;
        if (!w.need(qsub, (size_t)nfail * dim * 4) || !w.need(dsub, (size_t)nfail * k * 4) ||
            !w.need(isub, (size_t)nfail * k * 4)) {
            set_error("Memory allocation error");
            return false;
        }
        KNN_TRY(idx_gather_columns(query, ldq, dim, w.fb_list.as<int>(), nfail, qsub.as<float>(), s),
                "gather");
        knnhip_options po = o;
        po.force_fallback = 0;
        po.timing = 0;
        // the retry widens both the probe count and the candidate margin
        po.margin = std::max(2 * margin, kDefaultMargin);
        const int kp2 = std::max(k, std::min(k + po.margin, ix->R));
        const int P2 = std::min({nb, kMaxUnion / kp2, std::max(4 * P, kRetryProbes)});
        if (depth == 0 && P2 > P) {
            if (!index_search(ix, g_iw_retry, 1, qsub.as<float>(), nfail, nfail, k, P2, nullptr,
                              dsub.as<float>(), isub.as<int>(), po, s))
                return false;
        } else {
            const Plan pp = make_plan(ix->R, nfail, dim, k, po);
            if (!run_device(ix->ref, ix->ldr, ix->R, qsub.as<float>(), nfail, nfail, dim, k,
                            dsub.as<float>(), isub.as<int>(), pp, po, s))
                return false;
        }
        KNN_TRY(idx_scatter_results(dsub.as<float>(), isub.as<int>(), k, nfail, w.fb_list.as<int>(),
                                    Q, out_dist, out_idx, s),
                "scatter");
        KNN_TRY(hipStreamSynchronize(s), "sync");
    }
    ix->last_probes = P;
    ix->last_tasks = max_tasks;
    g_last_stream = s;
    for (float &m : g_stage_ms) m = -1.f;
    if (timing) {
        KNN_TRY(hipEventRecord(ev[5], s), "event");
        KNN_TRY(hipEventSynchronize(ev[5]), "event");
        ix->last_fail = nfail;
        for (int i = 0; i < 5; ++i) KNN_TRY(hipEventElapsedTime(&ix->last_ms[i], ev[i], ev[i + 1]), "event");
        KNN_TRY(hipEventElapsedTime(&ix->rank_ms[0], ev[0], evr[0]), "event");
        KNN_TRY(hipEventElapsedTime(&ix->rank_ms[1], evr[0], evr[1]), "event");
        KNN_TRY(hipEventElapsedTime(&ix->rank_ms[2], evr[1], ev[1]), "event");
        for (auto &e : ev) (void)hipEventDestroy(e);
        for (auto &e : evr) (void)hipEventDestroy(e);
        // stage slots: 0 ranking (norms, bounds, pick + tasks), 1 scan, 2 select, 3 fallback
        g_stage_ms[0] = ix->last_ms[0] + ix->last_ms[1];
        g_stage_ms[1] = ix->last_ms[2];
        g_stage_ms[2] = ix->last_ms[3];
        g_stage_ms[3] = ix->last_ms[4];
    }
    g_last_fb = w.fb_count.as<int>();
    return true;
}

}  // namespace
}  // namespace knnhip

// The opaque router handle of the C API: MLP dim -> H -> H -> nb.
struct knnhip_router {
    int dim = 0, H = 0, nb = 0;
    knnhip::RouterLayout lay{};
    knnhip::DevBuf params, grad, m, v;
    int step = 0;
    // batch workspace (grown on demand)
    knnhip::DevBuf h1, h2, z3, dz3, d2, d1, loss, lsum;
};

namespace knnhip {
namespace {

RouterLayout router_layout(int dim, int H, int nb) {
    RouterLayout l{};
    const int in[3] = {dim, H, H}, out[3] = {H, H, nb};
    long long off = 0;
    for (int i = 0; i < 3; ++i) {
        l.w_off[i] = off;
        l.w_len[i] = (long long)out[i] * in[i];
        off += l.w_len[i];
        l.b_off[i] = off;
        off += out[i];
        l.fan_in[i] = in[i];
    }
    l.total = off;
    return l;
}

// Batch buffers for n samples: the hidden layers, and with `train` the
// logits, the backward pass and the loss terms.
bool router_reserve(knnhip_router *r, int n, bool train) {
    auto need = [](DevBuf &b, size_t bytes) { return b.n >= bytes || b.alloc(bytes); };
    bool ok = need(r->h1, (size_t)n * r->H * 4) && need(r->h2, (size_t)n * r->H * 4);
    if (train)
        ok = ok && need(r->z3, (size_t)n * r->nb * 4) && need(r->dz3, (size_t)n * r->nb * 4) &&
             need(r->d2, (size_t)n * r->H * 4) && need(r->d1, (size_t)n * r->H * 4) &&
             need(r->loss, (size_t)n * 4) && need(r->lsum, 4);
    if (!ok) {
        set_error("Memory allocation error");
        return false;
    }
    return true;
}

GemmArgs gemm(const float *A, long long lda, bool ta, const float *B, long long ldb, bool tb,
              float *C, long long ldc, int M, int N, int K) {
    GemmArgs g{};
    g.A = A;
    g.lda = lda;
    g.ta = ta;
    g.B = B;
    g.ldb = ldb;
    g.tb = tb;
    g.C = C;
    g.ldc = ldc;
    g.M = M;
    g.N = N;
    g.K = K;
    return g;
}

// Hidden layers of n samples x (dim-major, leading dimension ld) into
// r->h1, r->h2 ([n][H]).
bool router_hidden(knnhip_router *r, const float *x, int n, long long ld, hipStream_t s) {
    const float *w = r->params.as<float>();
    const RouterLayout &L = r->lay;
    GemmArgs g = gemm(x, ld, true, w + L.w_off[0], r->dim, true, r->h1.as<float>(), r->H, n, r->H,
                      r->dim);
    g.bias = w + L.b_off[0];
    g.bias_mode = 1;
    g.relu = true;
    KNN_TRY(launch_gemm(g, s), "router layer 1");
    g = gemm(r->h1.as<float>(), r->H, false, w + L.w_off[1], r->H, true, r->h2.as<float>(), r->H, n,
             r->H, r->H);
    g.bias = w + L.b_off[1];
    g.bias_mode = 1;
    g.relu = true;
    KNN_TRY(launch_gemm(g, s), "router layer 2");
    return true;
}

// Forward + backward of a batch: loss terms in r->loss, gradients in
// r->grad (mean over the batch).  tgt[j*ldt + i]: block of neighbour j of
// sample i.
bool router_loss_grad(knnhip_router *r, const float *x, int n, long long ld, const int *tgt,
                      long long ldt, int k, hipStream_t s) {
    if (!router_reserve(r, n, true) || !router_hidden(r, x, n, ld, s)) return false;
    const float *w = r->params.as<float>();
    float *gr = r->grad.as<float>();
    const RouterLayout &L = r->lay;
    const int H = r->H, nb = r->nb;
    float *h1 = r->h1.as<float>(), *h2 = r->h2.as<float>(), *z3 = r->z3.as<float>(),
          *dz3 = r->dz3.as<float>(), *d2 = r->d2.as<float>(), *d1 = r->d1.as<float>();
    GemmArgs g = gemm(h2, H, false, w + L.w_off[2], H, true, z3, nb, n, nb, H);
    g.bias = w + L.b_off[2];
    g.bias_mode = 1;
    KNN_TRY(launch_gemm(g, s), "router layer 3");
    KNN_TRY(launch_softmax_ce(z3, n, nb, tgt, ldt, k, 1.0f / (float)n, dz3, r->loss.as<float>(), s),
            "router loss");
    // layer 3: gW3 = dZ3^T H2, gb3 = colsum dZ3, dZ2 = (dZ3 W3) * [H2 > 0]
    KNN_TRY(launch_gemm(gemm(dz3, nb, true, h2, H, false, gr + L.w_off[2], H, nb, H, n), s),
            "router grad W3");
    KNN_TRY(launch_colsum(dz3, n, nb, nb, gr + L.b_off[2], s), "router grad b3");
    g = gemm(dz3, nb, false, w + L.w_off[2], H, false, d2, H, n, H, nb);
    g.mask = h2;
    g.ldm = H;
    KNN_TRY(launch_gemm(g, s), "router back 3");
    // layer 2
    KNN_TRY(launch_gemm(gemm(d2, H, true, h1, H, false, gr + L.w_off[1], H, H, H, n), s),
            "router grad W2");
    KNN_TRY(launch_colsum(d2, n, H, H, gr + L.b_off[1], s), "router grad b2");
    g = gemm(d2, H, false, w + L.w_off[1], H, false, d1, H, n, H, H);
    g.mask = h1;
    g.ldm = H;
    KNN_TRY(launch_gemm(g, s), "router back 2");
    // layer 1: gW1 = dZ1^T X^T (x is dim-major)
    KNN_TRY(launch_gemm(gemm(d1, H, true, x, ld, true, gr + L.w_off[0], r->dim, H, r->dim, n), s),
            "router grad W1");
    KNN_TRY(launch_colsum(d1, n, H, H, gr + L.b_off[0], s), "router grad b1");
    return true;
}

constexpr float kAdamB1 = 0.9f, kAdamB2 = 0.999f, kAdamEps = 1e-8f;

bool router_step(knnhip_router *r, const float *x, int n, long long ld, const int *tgt,
                 long long ldt, int k, float lr, hipStream_t s) {
    if (!router_loss_grad(r, x, n, ld, tgt, ldt, k, s)) return false;
    r->step += 1;
    KNN_TRY(launch_adam(r->params.as<float>(), r->grad.as<float>(), r->m.as<float>(),
                        r->v.as<float>(), r->lay.total, lr, kAdamB1, kAdamB2, kAdamEps, r->step, s),
            "adam");
    return true;
}

// Block scores of Q queries (dim-major, ld) into scores[b*Q + q].
bool router_scores(knnhip_router *r, const float *q, int Q, long long ld, float *scores,
                   hipStream_t s) {
    if (!router_reserve(r, Q, false) || !router_hidden(r, q, Q, ld, s)) return false;
    const float *w = r->params.as<float>();
    GemmArgs g = gemm(w + r->lay.w_off[2], r->H, false, r->h2.as<float>(), r->H, true, scores, Q,
                      r->nb, Q, r->H);
    g.bias = w + r->lay.b_off[2];
    g.bias_mode = 2;
    KNN_TRY(launch_gemm(g, s), "router scores");
    return true;
}

// Training: exact neighbours of the training queries (plain engine), their
// blocks as targets, then `epochs` passes of shuffled mini-batches.
bool router_train(knnhip_router *r, const knnhip_index *ix, const float *tq, int n, long long ld,
                  int k, int epochs, int batch, float lr, unsigned long long seed, float *final_loss,
                  hipStream_t s) {
    DevBuf nd, ni, rb, tgt, perm, xb, tb, lsteps;
    const int steps_per_epoch = (n + batch - 1) / batch;
    if (!nd.alloc((size_t)k * n * 4) || !ni.alloc((size_t)k * n * 4) ||
        !rb.alloc((size_t)ix->R * 4) || !tgt.alloc((size_t)k * n * 4) ||
        !perm.alloc((size_t)epochs * n * 4) || !xb.alloc((size_t)r->dim * batch * 4) ||
        !tb.alloc((size_t)k * batch * 4) || !lsteps.alloc((size_t)steps_per_epoch * 4)) {
        set_error("Memory allocation error");
        return false;
    }
    knnhip_options po;
    memset(&po, 0, sizeof(po));
    if (!run_device(ix->ref, ix->ldr, ix->R, tq, n, ld, r->dim, k, nd.as<float>(), ni.as<int>(),
                    make_plan(ix->R, n, r->dim, k, po), po, s))
        return false;
    KNN_TRY(idx_ref_block(ix->block_start.as<int>(), ix->block_count.as<int>(), ix->nb,
                          ix->perm.as<int>(), rb.as<int>(), s),
            "block of each reference");
    KNN_TRY(launch_neighbour_blocks(ni.as<int>(), n, k, rb.as<int>(), tgt.as<int>(), s), "targets");
    std::vector<int> hp((size_t)epochs * n);
    unsigned long long st = seed;
    auto next = [&st]() {
        st += 0x9E3779B97F4A7C15ull;
        unsigned long long z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    for (int e = 0; e < epochs; ++e) {
        int *p = hp.data() + (size_t)e * n;
        for (int i = 0; i < n; ++i) p[i] = i;
        for (int i = n - 1; i > 0; --i) std::swap(p[i], p[next() % (unsigned long long)(i + 1)]);
    }
    KNN_TRY(hipMemcpyAsync(perm.p, hp.data(), hp.size() * 4, hipMemcpyHostToDevice, s), "copy");
    for (int e = 0; e < epochs; ++e)
        for (int b = 0; b < steps_per_epoch; ++b) {
            const int off = b * batch, nbat = std::min(batch, n - off);
            const int *list = perm.as<int>() + (size_t)e * n + off;
            KNN_TRY(idx_gather_columns(tq, ld, r->dim, list, nbat, xb.as<float>(), s), "gather");
            KNN_TRY(launch_gather_int_columns(tgt.as<int>(), n, k, list, nbat, tb.as<int>(), s),
                    "gather");
            if (!router_step(r, xb.as<float>(), nbat, nbat, tb.as<int>(), nbat, k, lr, s)) return false;
            KNN_TRY(launch_sum(r->loss.as<float>(), nbat, lsteps.as<float>() + b, s), "loss");
        }
    std::vector<float> ls(steps_per_epoch);
    KNN_TRY(hipMemcpyAsync(ls.data(), lsteps.p, ls.size() * 4, hipMemcpyDeviceToHost, s), "copy");
    KNN_TRY(hipStreamSynchronize(s), "sync");
    double tot = 0.0;
    for (float v : ls) tot += v;
    if (final_loss) *final_loss = (float)(tot / n);
    return true;
}

}  // namespace
}  // namespace knnhip

namespace knnhip {
namespace {

// ---------------------------------------------------------------------------
// Batched search: many independent problems, one scan launch per (dim, k)
// class.  Within a class the problems' reference sets are laid out one after
// another (each starting on a 32-point tile) in one packed copy, and the
// task-mode scan runs tasks of (one slice of one problem's references, up
// to 32 of that problem's queries) for every problem at once.  The rows-mode
// re-rank maps positions back to each problem's own indexes, so every query
// gets exactly the result of a search of its own problem.  The slice length
// (references per task) is the class's tile shape, picked by the autotuner
// or, untuned, by the cost model below.
// ---------------------------------------------------------------------------
struct BatchWork {
    DevBuf packed, pm, norms, perm, rmax, starts, counts, query, qpm, qnorm, theta, tasks, qlist,
        rows, pval, pidx, dist, idx, fb_count, fb_list, qsub, dsub, isub, flist;
    bool need(DevBuf &b, size_t bytes) { return b.n >= bytes || b.alloc(bytes); }
};
BatchWork g_bw;

constexpr int kBatchSlices[] = {2048, 8192, 32768, 131072, 1 << 20};

struct BatchTuneKey {
    int dim, k;
    bool operator<(const BatchTuneKey &o) const { return dim != o.dim ? dim < o.dim : k < o.k; }
};
std::map<BatchTuneKey, int> g_batch_table;  // (dim, k) -> slice length

struct BatchClass {
    int dim = 0, k = 0;
    std::vector<const knnhip_problem *> probs;
};

// Tasks of a class for slice length L: per problem ceil(R/L) slices times
// ceil(Q/32) query groups.
long long batch_tasks(const BatchClass &c, int L) {
    long long t = 0;
    for (const knnhip_problem *p : c.probs)
        t += (long long)((p->ref_nb + L - 1) / L) * ((p->query_nb + kQueriesPerWave - 1) / kQueriesPerWave);
    return t;
}

int batch_max_slices(const BatchClass &c, int L) {
    int m = 1;
    for (const knnhip_problem *p : c.probs) m = std::max(m, (p->ref_nb + L - 1) / L);
    return m;
}

// Untuned choice: the slice lengths whose per-query row count fits the
// re-rank's union (rows * kp <= kMaxUnion), costed as rounds of resident
// blocks times (steps per wave + a fixed per-task cost of ~3 steps).
int batch_default_slice(const BatchClass &c, int kp, int resident) {
    int best = 0;
    double best_cost = 1e300;
    for (int L : kBatchSlices) {
        if ((long long)batch_max_slices(c, L) * kp > kMaxUnion) continue;
        double cost = 0.0;
        // per-task work: a task's four waves share its slice's 128-ref steps
        long long tasks = 0, steps = 0;
        for (const knnhip_problem *p : c.probs) {
            const long long ns = (p->ref_nb + L - 1) / L, ng = (p->query_nb + 31) / 32;
            tasks += ns * ng;
            const long long span = std::min<long long>(L, p->ref_nb);
            steps += ns * ng * ((span + kRefsPerStep - 1) / kRefsPerStep);
        }
        const double avg = (double)steps / std::max<long long>(1, tasks) / 4.0 + 3.0;
        cost = std::ceil((double)tasks / resident) * avg;
        if (cost < best_cost) {
            best_cost = cost;
            best = L;
        }
    }
    return best;
}

bool batch_class_run(const BatchClass &c, int L, const knnhip_options &o, hipStream_t s,
                     float *scan_ms) {
    BatchWork &w = g_bw;
    const int dim = c.dim, k = c.k;
    int maxR = 0;
    for (const knnhip_problem *p : c.probs) maxR = std::max(maxR, p->ref_nb);
    const Plan pl = make_plan(maxR, kQueriesPerWave, dim, k, o);
    if (!pl.use_scan) {
        // k too large for the scan's heaps: each problem on its own
        for (const knnhip_problem *p : c.probs)
            if (!run_device(p->ref, p->ref_ld, p->ref_nb, p->query, p->query_ld, p->query_nb, dim, k,
                            p->knn_dist, p->knn_index, make_plan(p->ref_nb, p->query_nb, dim, k, o), o,
                            s))
                return false;
        return true;
    }
    const int kp = pl.kp;
    const int np = (int)c.probs.size();
    const int pm_ld = (dim + 3) & ~3;
    // layout
    std::vector<int> starts(np), counts(np), qstart(np);
    long long n_pos = 0, Qtot = 0;
    for (int i = 0; i < np; ++i) {
        starts[i] = (int)n_pos;
        counts[i] = c.probs[i]->ref_nb;
        n_pos += (c.probs[i]->ref_nb + kRefsPerTile - 1) / kRefsPerTile * kRefsPerTile;
        qstart[i] = (int)Qtot;
        Qtot += c.probs[i]->query_nb;
    }
    if (n_pos > (1ll << 30) || Qtot > (1ll << 30)) {
        set_error("batched class too large");
        return false;
    }
    const int r_pad = (int)((n_pos + kRefsPerStep - 1) / kRefsPerStep * kRefsPerStep);
    const int P = batch_max_slices(c, L);
    if ((long long)P * kp > kMaxUnion) {
        set_error("slice length %d too short for k=%d", L, k);
        return false;
    }
    // tasks, query lists, and each query's rows (one per slice of its problem)
    std::vector<ScanTask> tasks;
    std::vector<int> qlist, rows((size_t)Qtot * P, -1);
    for (int i = 0; i < np; ++i) {
        const knnhip_problem *p = c.probs[i];
        const int ns = (p->ref_nb + L - 1) / L;
        for (int sl = 0; sl < ns; ++sl)
            for (int g = 0; g < p->query_nb; g += kQueriesPerWave) {
                const int t = (int)tasks.size();
                ScanTask tk;
                tk.ref_begin = starts[i] + sl * L;
                tk.ref_end = starts[i] + std::min(p->ref_nb, (sl + 1) * L);
                tk.qoff = t * kQueriesPerWave;
                tk.qcount = std::min(kQueriesPerWave, p->query_nb - g);
                tasks.push_back(tk);
                for (int j = 0; j < kQueriesPerWave; ++j) {
                    const int q = qstart[i] + std::min(g + j, p->query_nb - 1);
                    qlist.push_back(q);
                    if (j < tk.qcount) rows[(size_t)q * P + sl] = t * kQueriesPerWave + j;
                }
            }
    }
    const int ntasks = (int)tasks.size();
    const size_t nrow = (size_t)ntasks * kQueriesPerWave;
    if (!w.need(w.packed, (size_t)r_pad * pl.dpad * 4) || !w.need(w.pm, (size_t)n_pos * pm_ld * 4) ||
        !w.need(w.norms, (size_t)r_pad * 4) || !w.need(w.perm, (size_t)n_pos * 4) ||
        !w.need(w.rmax, 4) || !w.need(w.starts, (size_t)np * 4) || !w.need(w.counts, (size_t)np * 4) ||
        !w.need(w.query, (size_t)Qtot * dim * 4) || !w.need(w.qpm, (size_t)Qtot * pm_ld * 4) ||
        !w.need(w.qnorm, (size_t)Qtot * 4) || !w.need(w.theta, (size_t)Qtot * 4) ||
        !w.need(w.tasks, (size_t)ntasks * sizeof(ScanTask)) || !w.need(w.qlist, nrow * 4) ||
        !w.need(w.rows, (size_t)Qtot * P * 4) || !w.need(w.pval, nrow * kp * 4) ||
        !w.need(w.pidx, nrow * kp * 4) || !w.need(w.dist, (size_t)Qtot * k * 4) ||
        !w.need(w.idx, (size_t)Qtot * k * 4) || !w.need(w.fb_count, 4) ||
        !w.need(w.fb_list, (size_t)Qtot * 4)) {
        set_error("Memory allocation error");
        return false;
    }
    KNN_TRY(hipMemcpyAsync(w.starts.p, starts.data(), (size_t)np * 4, hipMemcpyHostToDevice, s), "copy");
    KNN_TRY(hipMemcpyAsync(w.counts.p, counts.data(), (size_t)np * 4, hipMemcpyHostToDevice, s), "copy");
    KNN_TRY(hipMemcpyAsync(w.tasks.p, tasks.data(), (size_t)ntasks * sizeof(ScanTask),
                           hipMemcpyHostToDevice, s),
            "copy");
    KNN_TRY(hipMemcpyAsync(w.qlist.p, qlist.data(), nrow * 4, hipMemcpyHostToDevice, s), "copy");
    KNN_TRY(hipMemcpyAsync(w.rows.p, rows.data(), rows.size() * 4, hipMemcpyHostToDevice, s), "copy");
    // references: per problem norms + packed tiles + point-major rows at its start
    KNN_TRY(hipMemsetAsync(w.rmax.p, 0, 4, s), "memset");
    KNN_TRY(hipMemsetAsync(w.norms.p, 0, (size_t)r_pad * 4, s), "memset");
    if (r_pad > n_pos)
        KNN_TRY(hipMemsetAsync(w.packed.as<float>() + (size_t)n_pos * pl.dpad, 0,
                               (size_t)(r_pad - n_pos) * pl.dpad * 4, s),
                "memset");
    for (int i = 0; i < np; ++i) {
        const knnhip_problem *p = c.probs[i];
        NormArgs na{};
        na.p = p->ref;
        na.ld = p->ref_ld;
        na.n = p->ref_nb;
        na.n_pad = (p->ref_nb + kRefsPerTile - 1) / kRefsPerTile * kRefsPerTile;
        na.dim = dim;
        na.oct = pl.oct;
        na.out = w.norms.as<float>() + starts[i];
        na.maxbits = w.rmax.as<unsigned>();
        na.pm = w.pm.as<float>() + (size_t)starts[i] * pm_ld;
        na.pm_ld = pm_ld;
        na.packed = w.packed.as<float>() + (size_t)starts[i] * pl.dpad;
        KNN_TRY(launch_norms(na, s), "norm kernel launch");
        KNN_TRY(hipMemcpy2DAsync(w.query.as<float>() + qstart[i], (size_t)Qtot * 4, p->query,
                                 (size_t)p->query_ld * 4, (size_t)p->query_nb * 4, dim,
                                 hipMemcpyDeviceToDevice, s),
                "query copy");
    }
    KNN_TRY(idx_batch_perm(w.starts.as<int>(), w.counts.as<int>(), np, (int)n_pos, w.perm.as<int>(), s),
            "positions");
    NormArgs nq{};
    nq.p = w.query.as<float>();
    nq.ld = Qtot;
    nq.n = (int)Qtot;
    nq.n_pad = (int)Qtot;
    nq.dim = dim;
    nq.oct = (dim + 7) / 8;
    nq.out = w.qnorm.as<float>();
    nq.pm = w.qpm.as<float>();
    nq.pm_ld = pm_ld;
    KNN_TRY(launch_norms(nq, s), "norm kernel launch");
    KNN_TRY(hipMemsetD32Async((hipDeviceptr_t)w.theta.p, 0xFF800000u, Qtot, s), "memset");
    KNN_TRY(hipMemsetAsync(w.fb_count.p, 0, 4, s), "memset");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (scan_ms) {
        KNN_TRY(hipEventCreate(&e0), "event");
        KNN_TRY(hipEventCreate(&e1), "event");
        KNN_TRY(hipEventRecord(e0, s), "event");
    }
    ScanArgs sa{};
    sa.packed = w.packed.as<float>();
    sa.oct = pl.oct;
    sa.R = (int)n_pos;
    sa.dim = dim;
    sa.query = w.query.as<float>();
    sa.ldq = Qtot;
    sa.Q = (int)Qtot;
    sa.rnorm = w.norms.as<float>();
    sa.theta = w.theta.as<unsigned>();
    sa.slice_len = r_pad;
    sa.slice_span = (long long)r_pad * pl.dpad * 4;
    sa.r_pad = r_pad;
    sa.nslices = 1;
    sa.kp = kp;
    sa.nqb = 1;
    sa.nchunks = pl.nchunks;
    sa.part_val = w.pval.as<float>();
    sa.part_idx = w.pidx.as<int>();
    sa.tasks = w.tasks.as<ScanTask>();
    sa.ntasks = ntasks;
    sa.qlist = w.qlist.as<int>();
    sa.query_pm = w.qpm.as<float>();
    sa.query_pm_ld = pm_ld;
    KNN_TRY(launch_scan(sa, pl.ks, pl.multi, 4, s), "scan kernel launch");
    if (scan_ms) KNN_TRY(hipEventRecord(e1, s), "event");
    FinalArgs fa{};
    fa.R = (int)n_pos;
    fa.query = w.query.as<float>();
    fa.ldq = Qtot;
    fa.Q = (int)Qtot;
    fa.dim = dim;
    fa.k = k;
    fa.kp = kp;
    fa.nrows = P;
    fa.rows = w.rows.as<int>();
    fa.perm = w.perm.as<int>();
    fa.maxc = std::min(kMaxCand, std::max(2 * kp, 256));
    fa.part_val = w.pval.as<float>();
    fa.part_idx = w.pidx.as<int>();
    fa.ref_pm = w.pm.as<float>();
    fa.query_pm = w.qpm.as<float>();
    fa.pm_ld = pm_ld;
    fa.qnorm = w.qnorm.as<float>();
    fa.rmax = w.rmax.as<unsigned>();
    fa.theta = w.theta.as<unsigned>();
    fa.force_fallback = o.force_fallback;
    fa.out_dist = w.dist.as<float>();
    fa.out_idx = w.idx.as<int>();
    fa.fb_count = w.fb_count.as<int>();
    fa.fb_list = w.fb_list.as<int>();
    KNN_TRY(launch_select(fa, s), "select kernel launch");
    // certificate failures: the plain search of their own problem
    int nfail = 0;
    KNN_TRY(hipMemcpyAsync(&nfail, w.fb_count.p, 4, hipMemcpyDeviceToHost, s), "copy");
    KNN_TRY(hipStreamSynchronize(s), "sync");
    if (scan_ms) {
        KNN_TRY(hipEventElapsedTime(scan_ms, e0, e1), "event");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    g_batch_fail += nfail;
    if (nfail > 0) {
        std::vector<int> fl(nfail);
        KNN_TRY(hipMemcpy(fl.data(), w.fb_list.p, (size_t)nfail * 4, hipMemcpyDeviceToHost), "copy");
        std::sort(fl.begin(), fl.end());
        size_t a = 0;
        for (int i = 0; i < np && a < fl.size(); ++i) {
            const knnhip_problem *p = c.probs[i];
            size_t b = a;
            while (b < fl.size() && fl[b] < qstart[i] + p->query_nb) ++b;
            const int n = (int)(b - a);
            if (n == 0) continue;
            if (!w.need(w.flist, (size_t)n * 4) || !w.need(w.qsub, (size_t)n * dim * 4) ||
                !w.need(w.dsub, (size_t)n * k * 4) || !w.need(w.isub, (size_t)n * k * 4)) {
                set_error("Memory allocation error");
                return false;
            }
            KNN_TRY(hipMemcpy(w.flist.p, fl.data() + a, (size_t)n * 4, hipMemcpyHostToDevice), "copy");
            KNN_TRY(idx_gather_columns(w.query.as<float>(), Qtot, dim, w.flist.as<int>(), n,
                                       w.qsub.as<float>(), s),
                    "gather");
            knnhip_options o2 = o;
            o2.force_fallback = 0;
            o2.timing = 0;
            if (!run_device(p->ref, p->ref_ld, p->ref_nb, w.qsub.as<float>(), n, n, dim, k,
                            w.dsub.as<float>(), w.isub.as<int>(), make_plan(p->ref_nb, n, dim, k, o2),
                            o2, s))
                return false;
            KNN_TRY(idx_scatter_results(w.dsub.as<float>(), w.isub.as<int>(), k, n, w.flist.as<int>(),
                                        (int)Qtot, w.dist.as<float>(), w.idx.as<int>(), s),
                    "scatter");
            KNN_TRY(hipStreamSynchronize(s), "sync");
            a = b;
        }
    }
    // each problem's k x query_nb outputs
    for (int i = 0; i < np; ++i) {
        const knnhip_problem *p = c.probs[i];
        KNN_TRY(hipMemcpy2DAsync(p->knn_dist, (size_t)p->query_nb * 4, w.dist.as<float>() + qstart[i],
                                 (size_t)Qtot * 4, (size_t)p->query_nb * 4, k, hipMemcpyDeviceToDevice, s),
                "output copy");
        KNN_TRY(hipMemcpy2DAsync(p->knn_index, (size_t)p->query_nb * 4, w.idx.as<int>() + qstart[i],
                                 (size_t)Qtot * 4, (size_t)p->query_nb * 4, k, hipMemcpyDeviceToDevice, s),
                "output copy");
    }
    return true;
}

std::vector<BatchClass> batch_classes(const knnhip_problem *probs, int n) {
    std::map<BatchTuneKey, int> at;
    std::vector<BatchClass> out;
    for (int i = 0; i < n; ++i) {
        const BatchTuneKey key{probs[i].dim, probs[i].k};
        auto it = at.find(key);
        if (it == at.end()) {
            it = at.emplace(key, (int)out.size()).first;
            out.emplace_back();
            out.back().dim = key.dim;
            out.back().k = key.k;
        }
        out[it->second].probs.push_back(&probs[i]);
    }
    return out;
}

bool batch_args(const knnhip_problem *probs, int n) {
    if (!probs || n < 1) {
        set_error("bad batched arguments");
        return false;
    }
    for (int i = 0; i < n; ++i) {
        const knnhip_problem &p = probs[i];
        if (!check_args(p.ref, p.ref_nb, p.query, p.query_nb, p.dim, p.k, p.knn_dist, p.knn_index) ||
            p.ref_ld < p.ref_nb || p.query_ld < p.query_nb) {
            set_error("problem %d: bad arguments", i);
            return false;
        }
    }
    return true;
}

int batch_resident() {
    // two resident scan blocks per CU is the common case
    return 2 * std::max(1, g_cu_count);
}

}  // namespace
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
    const Plan p = make_plan(ref_nb, query_nb, dim, k, o);
    g_last_plan = plan_json(p, ref_nb, query_nb, dim, k);
    return run_device(ref, ref_ld, ref_nb, query, query_ld, query_nb, dim, k, knn_dist, knn_index,
                      p, o, s)
               ? 1
               : 0;
}

int knnhip_last_fallback_count(void) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_last_fb) return 0;
    int n = -1;
    if (hipStreamSynchronize(g_last_stream) != hipSuccess ||
        hipMemcpy(&n, g_last_fb, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("cannot read the fallback count");
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

int knnhip_index_build(const float *ref, int ref_nb, long long ref_ld, int dim, int nblocks,
                       int iters, knnhip_index **out, void *stream) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!out || !ref || ref_nb < 1 || dim < 1 || ref_ld < ref_nb || nblocks < 1 ||
        nblocks > ref_nb || nblocks > knnhip::kIndexMaxBlocks || iters < 0) {
        set_error("bad index arguments");
        return 0;
    }
    *out = nullptr;
    if (!ensure_device()) return 0;
    knnhip_index *ix = new knnhip_index();
    if (!index_build(ix, ref, ref_nb, ref_ld, dim, nblocks, iters, (hipStream_t)stream)) {
        delete ix;
        return 0;
    }
    *out = ix;
    return 1;
}

int knnhip_index_search(knnhip_index *ix, const float *query, int query_nb, long long query_ld,
                        int k, int probes, const float *scores, float *knn_dist, int *knn_index,
                        const knnhip_options *opt, void *stream) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!ix) {
        set_error("null index");
        return 0;
    }
    if (!check_args(ix->ref, ix->R, query, query_nb, ix->dim, k, knn_dist, knn_index)) return 0;
    if (query_ld < query_nb || probes < 1) {
        set_error("bad search arguments");
        return 0;
    }
    if (!ensure_device()) return 0;
    knnhip_options o = opt ? *opt : default_options(0);
    char buf[256];
    snprintf(buf, sizeof(buf),
             "{\"index_blocks\": %d, \"probes\": %d, \"router\": %s, \"query_nb\": %d, \"k\": %d}",
             ix->nb, std::min(probes, ix->nb), scores ? "true" : "false", query_nb, k);
    g_last_plan = buf;
    return index_search(ix, g_iw, 0, query, query_nb, query_ld, k, probes, scores, knn_dist,
                        knn_index, o, (hipStream_t)stream)
               ? 1
               : 0;
}

void knnhip_index_free(knnhip_index *ix) {
    std::lock_guard<std::mutex> lock(g_mu);
    delete ix;
}

int knnhip_index_save(const knnhip_index *ix, const char *path) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!ix || !path) {
        set_error("bad index save arguments");
        return 0;
    }
    return index_save(ix, path) ? 1 : 0;
}

int knnhip_index_load(const char *path, const float *ref, int ref_nb, long long ref_ld, int dim,
                      knnhip_index **out, void *stream) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!path || !ref || !out || ref_nb < 1 || dim < 1 || ref_ld < ref_nb) {
        set_error("bad index load arguments");
        return 0;
    }
    *out = nullptr;
    if (!ensure_device()) return 0;
    knnhip_index *ix = new knnhip_index();
    if (!index_load(ix, path, ref, ref_nb, ref_ld, dim, (hipStream_t)stream)) {
        delete ix;
        return 0;
    }
    *out = ix;
    return 1;
}

int knnhip_index_stage_ms(const knnhip_index *ix, float *ms, int n) {
    if (!ix || !ms) return 0;
    const float all[7] = {ix->rank_ms[0], ix->rank_ms[1], ix->rank_ms[2], ix->last_ms[1],
                          ix->last_ms[2], ix->last_ms[3], ix->last_ms[4]};
    for (int i = 0; i < n && i < 7; ++i) ms[i] = all[i];
    return 1;
}

int knnhip_index_info(const knnhip_index *ix, int *nblocks, int *n_pos, float *build_ms) {
    if (!ix) return 0;
    if (nblocks) *nblocks = ix->nb;
    if (n_pos) *n_pos = ix->n_pos;
    if (build_ms) *build_ms = ix->build_ms;
    return 1;
}

int knnhip_index_layout(const knnhip_index *ix, int *block_start, int *block_count, int *ref_block,
                        void *stream) {
    if (!ix) {
        set_error("null index");
        return 0;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!ensure_device()) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (ref_block)
        KNN_TRY(idx_ref_block(ix->block_start.as<int>(), ix->block_count.as<int>(), ix->nb,
                                ix->perm.as<int>(), ref_block, s),
                  "block of each reference");
    if (block_start)
        KNN_TRY(hipMemcpyAsync(block_start, ix->block_start.p, (size_t)ix->nb * 4,
                                 hipMemcpyDeviceToHost, s),
                  "copy");
    if (block_count)
        KNN_TRY(hipMemcpyAsync(block_count, ix->block_count.p, (size_t)ix->nb * 4,
                                 hipMemcpyDeviceToHost, s),
                  "copy");
    if (block_start || block_count) KNN_TRY(hipStreamSynchronize(s), "sync");
    return 1;
}

int knnhip_router_create(int dim, int hidden, int nblocks, unsigned long long seed,
                         knnhip_router **out) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!out || dim < 1 || hidden < 1 || nblocks < 1 || dim > (1 << 16) || hidden > (1 << 16) ||
        nblocks > kIndexMaxBlocks) {
        set_error("bad router arguments");
        return 0;
    }
    *out = nullptr;
    if (!ensure_device()) return 0;
    knnhip_router *r = new knnhip_router();
    r->dim = dim;
    r->H = hidden;
    r->nb = nblocks;
    r->lay = router_layout(dim, hidden, nblocks);
    const size_t bytes = (size_t)r->lay.total * 4;
    bool ok = r->params.alloc(bytes) && r->grad.alloc(bytes) && r->m.alloc(bytes) && r->v.alloc(bytes);
    ok = ok && hipMemset(r->m.p, 0, bytes) == hipSuccess && hipMemset(r->v.p, 0, bytes) == hipSuccess &&
         launch_init_params(r->params.as<float>(), r->lay, seed, nullptr) == hipSuccess &&
         hipDeviceSynchronize() == hipSuccess;
    if (!ok) {
        if (g_last_error.empty()) set_error("router allocation or init failed");
        delete r;
        return 0;
    }
    *out = r;
    return 1;
}

void knnhip_router_free(knnhip_router *r) {
    std::lock_guard<std::mutex> lock(g_mu);
    delete r;
}

// On-disk format of a router: magic "KNNHIPRT", uint32 version (1), int32 dim,
// hidden, nblocks, int64 parameter count, then the FP32 parameter vector
// (the layout of knnhip_router_params).  Optimiser state is not stored.
int knnhip_router_save(const knnhip_router *r, const char *path) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!r || !path) {
        set_error("bad router save arguments");
        return 0;
    }
    FILE *f = fopen(path, "wb");
    if (!f) {
        set_error("cannot open %s for writing", path);
        return 0;
    }
    static const char magic[8] = {'K', 'N', 'N', 'H', 'I', 'P', 'R', 'T'};
    const unsigned version = 1;
    const int hdr[3] = {r->dim, r->H, r->nb};
    const long long n = r->lay.total;
    FileOut o{f};
    o.put(magic, 8);
    o.put(&version, 4);
    o.put(hdr, sizeof(hdr));
    o.put(&n, 8);
    o.put_dev(r->params, (size_t)n * 4);
    const bool ok = o.ok && fclose(f) == 0;
    if (!ok) set_error("cannot write %s", path);
    return ok ? 1 : 0;
}

int knnhip_router_load(const char *path, knnhip_router **out) {
    if (!path || !out) {
        std::lock_guard<std::mutex> lock(g_mu);
        set_error("bad router load arguments");
        return 0;
    }
    FILE *f = fopen(path, "rb");
    if (!f) {
        std::lock_guard<std::mutex> lock(g_mu);
        set_error("cannot open %s", path);
        return 0;
    }
    char magic[8] = {0};
    unsigned version = 0;
    int hdr[3] = {0};
    long long n = 0;
    FileIn in{f};
    in.get(magic, 8);
    in.get(&version, 4);
    in.get(hdr, sizeof(hdr));
    in.get(&n, 8);
    if (!in.ok || memcmp(magic, "KNNHIPRT", 8) != 0 || version != 1 || n != router_layout(hdr[0], hdr[1], hdr[2]).total) {
        fclose(f);
        std::lock_guard<std::mutex> lock(g_mu);
        set_error("%s is not a router file of version 1", path);
        return 0;
    }
    std::vector<float> params((size_t)n);
    in.get(params.data(), (size_t)n * 4);
    char extra;
    const bool tail = fread(&extra, 1, 1, f) == 0;
    fclose(f);
    if (!in.ok || !tail) {
        std::lock_guard<std::mutex> lock(g_mu);
        set_error("%s is truncated or has trailing data", path);
        return 0;
    }
    knnhip_router *r = nullptr;
    if (!knnhip_router_create(hdr[0], hdr[1], hdr[2], 0, &r)) return 0;
    if (!knnhip_router_params(r, params.data(), n, 1)) {
        knnhip_router_free(r);
        return 0;
    }
    *out = r;
    return 1;
}

long long knnhip_router_num_params(const knnhip_router *r) { return r ? r->lay.total : 0; }

int knnhip_router_params(knnhip_router *r, float *host, long long n, int set) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!r || !host || n != r->lay.total) {
        set_error("bad router parameter copy");
        return 0;
    }
    const hipError_t e = set ? hipMemcpy(r->params.p, host, (size_t)n * 4, hipMemcpyHostToDevice)
                             : hipMemcpy(host, r->params.p, (size_t)n * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        set_error("router parameter copy: %s", hipGetErrorString(e));
        return 0;
    }
    if (set) {
        r->step = 0;
        if (hipMemset(r->m.p, 0, (size_t)n * 4) != hipSuccess ||
            hipMemset(r->v.p, 0, (size_t)n * 4) != hipSuccess) {
            set_error("router optimiser reset failed");
            return 0;
        }
    }
    return 1;
}

static bool router_batch_args(const knnhip_router *r, const float *x, int n, long long ld,
                              const int *tgt, long long ldt, int k) {
    if (!r || !x || !tgt || n < 1 || ld < n || ldt < n || k < 1 || k > kRouterMaxK) {
        set_error("bad router batch arguments");
        return false;
    }
    return true;
}

int knnhip_router_loss_grad(knnhip_router *r, const float *x, int n, long long ld, const int *tgt,
                            long long ldt, int k, float *grad, float *loss, void *stream) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!router_batch_args(r, x, n, ld, tgt, ldt, k)) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (!router_loss_grad(r, x, n, ld, tgt, ldt, k, s)) return 0;
    if (grad)
        KNN_TRY(hipMemcpyAsync(grad, r->grad.p, (size_t)r->lay.total * 4, hipMemcpyDeviceToDevice, s),
                "copy");
    if (loss) {
        KNN_TRY(launch_sum(r->loss.as<float>(), n, r->lsum.as<float>(), s), "loss");
        float v = 0.0f;
        KNN_TRY(hipMemcpyAsync(&v, r->lsum.p, 4, hipMemcpyDeviceToHost, s), "copy");
        KNN_TRY(hipStreamSynchronize(s), "sync");
        *loss = v / (float)n;
    }
    return 1;
}

int knnhip_router_step(knnhip_router *r, const float *x, int n, long long ld, const int *tgt,
                       long long ldt, int k, float lr, void *stream) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!router_batch_args(r, x, n, ld, tgt, ldt, k)) return 0;
    return router_step(r, x, n, ld, tgt, ldt, k, lr, (hipStream_t)stream) ? 1 : 0;
}

int knnhip_router_train(knnhip_router *r, const knnhip_index *ix, const float *train_query, int n,
                        long long ld, int k, int epochs, int batch, float lr,
                        unsigned long long seed, float *final_loss, void *stream) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!r || !ix || !train_query || n < 1 || ld < n || k < 1 || k > kRouterMaxK || k > ix->R ||
        epochs < 1 || batch < 1 || r->nb != ix->nb || r->dim != ix->dim) {
        set_error("bad router training arguments");
        return 0;
    }
    if (!ensure_device()) return 0;
    return router_train(r, ix, train_query, n, ld, k, epochs, std::min(batch, n), lr, seed,
                        final_loss, (hipStream_t)stream)
               ? 1
               : 0;
}

int knnhip_router_scores(knnhip_router *r, const float *query, int query_nb, long long query_ld,
                         float *scores, void *stream) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!r || !query || !scores || query_nb < 1 || query_ld < query_nb) {
        set_error("bad router score arguments");
        return 0;
    }
    return router_scores(r, query, query_nb, query_ld, scores, (hipStream_t)stream) ? 1 : 0;
}

int knnhip_search_batched(const knnhip_problem *problems, int count, const knnhip_options *opt,
                          void *stream) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!batch_args(problems, count) || !ensure_device()) return 0;
    const knnhip_options o = opt ? *opt : default_options(0);
    hipStream_t s = (hipStream_t)stream;
    g_batch_fail = 0;
    std::string plan = "{\"classes\": [";
    bool first = true;
    for (const BatchClass &c : batch_classes(problems, count)) {
        const Plan pl = make_plan(1 << 20, kQueriesPerWave, c.dim, c.k, o);
        auto it = g_batch_table.find(BatchTuneKey{c.dim, c.k});
        int L = it != g_batch_table.end() ? it->second : batch_default_slice(c, std::max(pl.kp, c.k), batch_resident());
        if (L == 0) L = kBatchSlices[sizeof(kBatchSlices) / sizeof(int) - 1];
        char buf[160];
        snprintf(buf, sizeof(buf), "%s{\"dim\": %d, \"k\": %d, \"problems\": %d, \"slice\": %d, \"tuned\": %s}",
                 first ? "" : ", ", c.dim, c.k, (int)c.probs.size(), L,
                 it != g_batch_table.end() ? "true" : "false");
        plan += buf;
        first = false;
        if (!batch_class_run(c, L, o, s, nullptr)) return 0;
    }
    g_last_plan = plan + "]}";
    g_last_fb = nullptr;
    return 1;
}

long long knnhip_batched_last_failures(void) { return g_batch_fail; }

int knnhip_batched_autotune(const knnhip_problem *problems, int count, int reps, void *stream) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!batch_args(problems, count) || !ensure_device()) return 0;
    knnhip_options o = default_options(0);
    hipStream_t s = (hipStream_t)stream;
    for (const BatchClass &c : batch_classes(problems, count)) {
        const Plan pl = make_plan(1 << 20, kQueriesPerWave, c.dim, c.k, o);
        const int kp = std::max(pl.kp, c.k);
        int best = 0;
        float best_ms = 0.f;
        for (int L : kBatchSlices) {
            if ((long long)batch_max_slices(c, L) * kp > kMaxUnion) continue;
            if (batch_tasks(c, L) > (1ll << 24)) continue;
            float tot = 0.f;
            for (int r = 0; r < std::max(1, reps); ++r) {
                float ms = 0.f;
                if (!batch_class_run(c, L, o, s, &ms)) return 0;
                tot += ms;
            }
            if (!best || tot < best_ms) {
                best = L;
                best_ms = tot;
            }
        }
        if (best) g_batch_table[BatchTuneKey{c.dim, c.k}] = best;
    }
    return 1;
}

int knnhip_batched_table(char *buf, int len) {
    std::lock_guard<std::mutex> lock(g_mu);
    std::string t = "[";
    bool first = true;
    for (const auto &kv : g_batch_table) {
        char e[96];
        snprintf(e, sizeof(e), "%s{\"dim\": %d, \"k\": %d, \"slice\": %d}", first ? "" : ", ",
                 kv.first.dim, kv.first.k, kv.second);
        t += e;
        first = false;
    }
    t += "]";
    if (buf && len > 0) {
        snprintf(buf, (size_t)len, "%s", t.c_str());
    }
    return (int)t.size() + 1;
}

int knnhip_batched_set_slice(int dim, int k, int slice_len) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (dim < 1 || k < 1 || slice_len < 0 || slice_len % kRefsPerStep) {
        set_error("bad tile-table entry");
        return 0;
    }
    if (slice_len == 0)
        g_batch_table.erase(BatchTuneKey{dim, k});
    else
        g_batch_table[BatchTuneKey{dim, k}] = slice_len;
    return 1;
}

const char *knnhip_version(void) { return "knnhip 0.2 (gfx950)"; }

}  // extern "C"
