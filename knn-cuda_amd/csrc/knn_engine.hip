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
#include <mutex>
#include <string>

#include "../../include/knncuda.h"
#include "knn_common.h"

namespace knnhip {

// Launchers defined in knn_scan.hip / knn_select.hip.
int scan_pick_ks(int dim);
size_t scan_lds_bytes(int kp, int waves);
int scan_blocks_per_cu(int ks, bool multi, int waves, int kp);
hipError_t launch_scan(const ScanArgs &a, int ks, bool multi, int waves, hipStream_t s);
hipError_t launch_norms(const NormArgs &a, hipStream_t s);
hipError_t launch_select(const FinalArgs &a, hipStream_t s);
hipError_t launch_exact(const ExactArgs &a, int grid, hipStream_t s);
hipError_t launch_gen_uniform(float *out, int n, int dim, uint64_t seed, float scale,
                              hipStream_t s);
hipError_t launch_gen_mixture(float *out, int n, int dim, int ncomp, float sigma,
                              uint64_t centre_seed, uint64_t point_seed, hipStream_t s);

namespace {

constexpr int kDefaultMargin = 16;
constexpr int kExactGrid = 512;
constexpr size_t kScanLdsMax = 160 * 1024;

std::mutex g_mu;
std::string g_last_error;
float g_stage_ms[4] = {-1.f, -1.f, -1.f, -1.f};
hipStream_t g_last_stream = nullptr;
std::string g_last_plan;
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
    kFbKey, kFbIdx, kOutDist, kOutIdx, kRefPm, kQryPm, kPacked, kTheta, kNumSlots
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
};

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
    const int margin = o.margin > 0 ? o.margin : kDefaultMargin;
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
    return p;
}

std::string plan_json(const Plan &p, int R, int Q, int dim, int k) {
    char buf[512];
    snprintf(buf, sizeof(buf),
             "{\"ref_nb\": %d, \"query_nb\": %d, \"dim\": %d, \"k\": %d, \"scan\": %s, "
             "\"ks\": %d, \"chunks\": %d, \"dpad\": %d, \"waves_per_block\": %d, \"kp\": %d, "
             "\"slices\": %d, \"slice_len\": %d, \"query_blocks\": %d, \"xcd_map\": %d, "
             "\"blocks_per_cu\": %d, \"cus\": %d, \"max_candidates\": %d}",
             R, Q, dim, k, p.use_scan ? "true" : "false", p.ks, p.nchunks, p.dpad, p.waves, p.kp,
             p.nslices, p.slice_len, p.nqb, p.xcd, p.blocks_per_cu, p.cus, p.maxc);
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
bool run_device(const float *ref, long long ldr, int R, const float *query, long long ldq, int Q,
                int dim, int k, float *out_dist, int *out_idx, const Plan &p,
                const knnhip_options &o, hipStream_t s) {
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
        float *rnorm = (float *)ws_get(kRNorm, (size_t)R * sizeof(float));
        float *qnorm = (float *)ws_get(kQNorm, (size_t)Q * sizeof(float));
        unsigned *rmax = (unsigned *)ws_get(kRMax, sizeof(unsigned));
        const size_t npart = (size_t)p.nslices * Q * p.kp;
        float *pval = (float *)ws_get(kPartVal, npart * sizeof(float));
        int *pidx = (int *)ws_get(kPartIdx, npart * sizeof(int));
        int *fb_list = (int *)ws_get(kFbList, (size_t)Q * sizeof(int));
        const int pm_ld = (dim + 3) & ~3;
        float *ref_pm = (float *)ws_get(kRefPm, (size_t)R * pm_ld * sizeof(float));
        float *qry_pm = (float *)ws_get(kQryPm, (size_t)Q * pm_ld * sizeof(float));
        float *packed = (float *)ws_get(kPacked, (size_t)p.r_pad * p.dpad * sizeof(float));
        unsigned *theta = (unsigned *)ws_get(kTheta, (size_t)Q * sizeof(unsigned));
        if (!rnorm || !qnorm || !rmax || !pval || !pidx || !fb_list || !ref_pm || !qry_pm ||
            !packed || !theta) {
            set_error("Memory allocation error");
            return false;
        }
        if (timing) KNN_TRY(hipEventRecord(g_ws.ev[0], s), "event");
        KNN_TRY(hipMemsetAsync(rmax, 0, sizeof(unsigned), s), "memset");
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
        na.packed = packed;
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
        sa.packed = packed;
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
        fa.nslices = p.nslices;
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

        // Queries that failed the certificate: the grid reads the count on
        // the device, so no host round trip is needed.
        ea.list = fb_list;
        ea.count = fb_count;
        KNN_TRY(launch_exact(ea, egrid, s), "exact scan launch");
        if (timing) KNN_TRY(hipEventRecord(g_ws.ev[4], s), "event");
    }
    g_last_stream = s;
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
bool run_host(const float *ref, int R, const float *query, int Q, int dim, int k,
              float *knn_dist, int *knn_index, int variant) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!check_args(ref, R, query, Q, dim, k, knn_dist, knn_index)) return false;
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
    if (!g_ws.ptr[kFbCount]) return 0;
    int n = -1;
    if (hipStreamSynchronize(g_last_stream) != hipSuccess ||
        hipMemcpy(&n, g_ws.ptr[kFbCount], sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
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

const char *knnhip_version(void) { return "knnhip 0.1 (gfx950)"; }

}  // extern "C"
