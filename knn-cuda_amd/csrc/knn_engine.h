// Internal interface of the host engine, shared by its modules:
//   knn_engine.hip      plan, workspace, the plain search and the drop-in API
//   knn_index_host.hip  block index build / search / save / load
//   knn_router_host.hip learned router training and scoring
//   knn_batch.hip       batched search and its slice-length autotuner
//   knn_stream.hip      searches streamed through a device memory budget
//   knn_hostio.hip      host transfers through pinned staging
// Every C entry point takes g_mu for its whole call, so the state below is
// touched by one host thread at a time.
#pragma once

#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/knncuda.h"
#include "knn_common.h"

namespace knnhip {

// Owned device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    // one owner per allocation: moves hand it over, copies do not exist
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) {
        o.p = nullptr;
        o.n = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
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

// Launchers defined in knn_scan.hip / knn_scan_bf16.hip / knn_select.hip.
int scan_pick_ks(int dim);
size_t scan_lds_bytes(int kp, int waves);
size_t scan_block_lds(int kp, int waves, int bmode, int oct);
int scan_blocks_per_cu(int ks, int bmode, int waves, int kp, int oct);
hipError_t launch_scan(const ScanArgs &a, int ks, int bmode, int waves, hipStream_t s);
// Batched scan (knn_scan.hip): one persistent launch over work items (class,
// task) of many classes, each with its own ScanArgs (device array cls) and
// instance variant (scan_multi_variant: packed layout of 2*ks dimensions, or
// nchunks chunks of 128).  next: a zeroed work counter.  lds: the largest
// class's scan_multi_lds(kp).
int scan_multi_variant(int dim, int *ks, int *nchunks);
size_t scan_multi_lds(int kp);
int scan_multi_blocks_per_cu(size_t lds);
hipError_t launch_scan_multi(const ScanArgs *cls, const int *variant, const int2 *work, int nwork, unsigned *next,
                             size_t lds, int grid, hipStream_t s);
int scan_bf3_nch(int dim);
size_t scan_bf3_lds_bytes(int kp);
hipError_t launch_scan_lds(const ScanArgs &a, int nch, bool bf3, hipStream_t s);
int scan_bf3_task_nch(int dim);
// tasks_per_group > 1 (with waves = 1): that many one-wave tasks per
// workgroup, each with its own heaps, so a block's consecutive query groups
// share a CU's caches.  merge (waves = 1): tasks of up to kMergeQueries
// queries keep sorted merged lists instead of heaps (kp <= kMergeMaxKp), and
// write their rows from their list offset.
hipError_t launch_scan_bf3_task(const ScanArgs &a, int nch, int waves, size_t lds, hipStream_t s,
                                int tasks_per_group = 1, bool merge = false);
hipError_t launch_pack_bf3_rows(const float *pm, int ld, int n_pad, int nch, unsigned short *out,
                                hipStream_t s);
// Residual form of the block-index scan (knn_scan_resid.hip): chunks for a
// dimension (0: too wide), its key error coefficients, the per-block
// centroids / residual tiles / radii of an index, and the task scan.
int scan_res_nch(int dim);
float res_eta_coeff(int dim, int nch);
float res_norm_up(int dim);
float res_norm_dn(int dim);
hipError_t launch_resid_build(const float *pm, int ld, int dim, int nch, const int *unit_start,
                              const int *unit_count, int nunits, float *cent, unsigned short *tiles, float *rnorm,
                              float *rho, hipStream_t s);
hipError_t launch_scan_res_task(const ScanArgs &a, int nch, int waves, size_t lds, hipStream_t s,
                                int tasks_per_group = 1, bool merge = false);
hipError_t launch_norms(const NormArgs &a, hipStream_t s);
hipError_t launch_norms_jobs(const NormArgs &a, const NormJob *jobs, const int2 *wg, int nwg,
                             hipStream_t s);
constexpr int kNormPoints = 64;  // points per workgroup of the norms pass
hipError_t launch_select(const FinalArgs &a, hipStream_t s);
hipError_t launch_exact(const ExactArgs &a, int grid, hipStream_t s);
hipError_t launch_gen_uniform(float *out, int n, int dim, uint64_t seed, float scale,
                              hipStream_t s);
hipError_t launch_gen_mixture(float *out, int n, int dim, int ncomp, float sigma,
                              uint64_t centre_seed, uint64_t point_seed, hipStream_t s);

// Candidates kept beyond k (kp = k + margin): the certificate needs the
// kp-th approximate key to clear the k-th exact distance by the FP32 error
// bound.  Large k goes with dense neighbourhoods, so it starts wider.
constexpr int kDefaultMargin = 16;
constexpr int kDefaultMarginLargeK = 64;
constexpr int kLargeK = 64;
// Certificate failures are retried through the scan with wider margins
// (kEscalationMargins[level - 1]) before the exact per-query scan.
constexpr int kEscalationMargins[] = {64, 256};
constexpr int kMaxEscalation = 2;
inline int default_margin(int k) { return k >= kLargeK ? kDefaultMarginLargeK : kDefaultMargin; }
constexpr int kExactGrid = 512;
constexpr size_t kScanLdsMax = 160 * 1024;

extern std::mutex g_mu;
extern std::string g_last_error;
extern float g_stage_ms[4];
extern hipStream_t g_last_stream;
extern const int *g_last_fb;  // device counter of the last search's exact-scan queries
extern std::string g_last_plan;
extern int g_cu_count;

void set_error(const char *fmt, ...);

// Grow-only device buffers, one per role.
enum Slot {
    kRefPad, kQryPad, kRNorm, kQNorm, kRMax, kPartVal, kPartIdx, kFbCount, kFbList,
    kFbKey, kFbIdx, kOutDist, kOutIdx, kRefPm, kQryPm, kPacked, kTheta, kPackedBf, kNonFinite,
    kNumSlots
};
struct Workspace {
    void *ptr[kNumSlots] = {};
    size_t cap[kNumSlots] = {};
    hipEvent_t ev[5] = {};
    bool have_events = false;
    hipEvent_t up_ev[8] = {};    // one per upload chunk
    hipEvent_t qry_ev = nullptr; // query copy on the upload stream
    hipStream_t copy = nullptr;  // upload stream
};
extern Workspace g_ws;
void *ws_get(Slot s, size_t bytes);

// While an Arena is active, ws_get carves each slot's buffer out of a
// caller-given region (256-byte aligned pieces) instead of the cached
// hipMalloc buffers, and run_device skips its wider-margin rescans (whose
// buffers it would allocate): queries that fail the certificate go straight
// to the exact per-query scan.
struct Arena {
    unsigned char *base = nullptr;
    size_t cap = 0, used = 0;
    void *slot[kNumSlots] = {};
    size_t size[kNumSlots] = {};
};
extern Arena *g_arena;

// Non-finite input of the plain search: one device word that the norms pass
// sets when a reference or query point has a NaN or infinite squared norm;
// the re-rank and the exact scan then write index -1 / distance NaN instead
// of ordering NaN keys.  run_device clears it at level 0, except while a
// NonfiniteHold is alive (a streamed search clears it once for all of its
// chunk searches).
struct NonfiniteHold {
    NonfiniteHold();
    ~NonfiniteHold();
};
extern const unsigned *g_last_nonfinite;  // flag of the last plain search
// Clears the flag on stream s (before a held sequence of searches).
bool nonfinite_clear(hipStream_t s);
// After the search's work on stream s: false with an error when the flag is
// set (the host entry points' check).
bool nonfinite_check(hipStream_t s);

// Launch plan of one call.
struct Plan {
    bool use_scan = false;  // false: exact full scan for every query
    int ks = 0;
    int bmode = kBRegs;       // where the scan keeps the -2q operands
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
inline int block_queries(const Plan &p) { return p.lds ? 128 : kQueriesPerWave; }

Plan make_plan(int R, int Q, int dim, int k, const knnhip_options &o);
// Bytes run_device takes from ws_get for this plan (each piece rounded up to
// 256 bytes): what an Arena must hold.
size_t run_device_bytes(const Plan &p, int R, int Q, int dim, int k);
std::string plan_json(const Plan &p, int R, int Q, int dim, int k);
bool check_args(const void *ref, int R, const void *query, int Q, int dim, int k, const void *dist,
                const void *idx);
bool ensure_device();

// Device-side order of consecutive calls.  Calls share the library's cached
// device buffers (the plain search's workspace, the index search's, the
// batched lanes') and a call may return while its kernels still run.  A
// call on another stream than the previous call's therefore first makes its
// stream wait for an event recorded at the end of the previous call, so no
// call's kernels touch a buffer that the previous call's kernels still use.
// Construct after ensure_device(), with g_mu held.
class CallOrder {
  public:
    explicit CallOrder(hipStream_t s);
    ~CallOrder();
    CallOrder(const CallOrder &) = delete;
    CallOrder &operator=(const CallOrder &) = delete;

  private:
    hipStream_t s_;
};
knnhip_options default_options(int variant);
// Key error coefficients of the bfloat16 three-term products over nch
// 16-dimension chunks (for FinalArgs::err_q / err_r).
void bf3_error_bound(int dim, int nch, double *err_q, double *err_r);
// Prints the last error on stdout as the reference's entry points do.
bool report(bool ok);

// Host-pointer references uploaded in chunks of whole slices on a copy
// stream; each chunk's norms, packing and slices are queued on the search
// stream as soon as its copy has landed.
struct Upload {
    const float *ref;    // host, dim rows of ld floats
    long long ld;
    hipStream_t stream;  // copy stream (non-blocking)
    int chunks;          // the plan's slices are grouped into this many chunks
};
constexpr int kUploadChunks = 8;
constexpr int kUploadMinRefs = 1 << 18;  // smaller sets are uploaded in one copy

// A search whose certificate-failure count is read back later: run_device
// queues an asynchronous copy of the count and an event instead of
// synchronising, and finish_device waits on the event and searches the
// failed queries again (queued on the same stream).
struct Deferred {
    bool pending = false;
    const float *ref = nullptr, *query = nullptr;
    long long ldr = 0, ldq = 0;
    int R = 0, Q = 0, dim = 0, k = 0;
    float *out_dist = nullptr;
    int *out_idx = nullptr;
    const int *fb_list = nullptr, *fb_count = nullptr;
    knnhip_options o{};
    int margin = 0, level = 0;
    hipStream_t s = nullptr;
    int *nfail_host = nullptr;  // pinned
    hipEvent_t ev = nullptr;
};
bool finish_device(Deferred &d);

// Search on device pointers: dim rows of ldr / ldq floats; results k x Q.
// level > 0: an escalation pass over queries that failed the certificate.
// up: the references are still on the host (ref is their device buffer,
// filled chunk by chunk); FP32 32-query plans only.
bool run_device(const float *ref, long long ldr, int R, const float *query, long long ldq, int Q,
                int dim, int k, float *out_dist, int *out_idx, const Plan &p,
                const knnhip_options &o, hipStream_t s, int level = 0, const Upload *up = nullptr,
                Deferred *dfr = nullptr);

// Copies between pageable host arrays (rows of width 4-byte words, pitch
// sld / dld words) and the device through pinned staging, ordered on stream
// s (knn_hostio.hip).  host_upload returns once the pieces are queued and
// the source is no longer read; host_download returns when dst holds the data.
bool host_upload(void *dst, long long dld, const void *src, long long sld, long long width, int rows,
                 hipStream_t s);
bool host_download(void *dst, long long dld, const void *src, long long sld, long long width, int rows,
                   hipStream_t s);

// Searches larger than the device memory they may use (knn_stream.hip).
// Device bytes of one host-pointer search of R references and Q queries,
// resident or streamed (two reference chunk buffers and merged results).
size_t search_footprint(int R, int Q, int dim, int k, const knnhip_options &o, bool streamed);
// Reference chunk and query group sizes of a streamed search within budget.
bool stream_chunks(int R, int Q, int dim, int k, const knnhip_options &o, size_t budget, int *rc,
                   int *qc);
// knnhip_set_device_budget's value, else 90 % of free plus cached memory.
size_t device_budget(bool *explicit_budget);
// Frees every cached workspace buffer (after the device is idle).
void ws_release();
bool run_streamed(const float *ref, int R, const float *query, int Q, int dim, int k,
                  float *knn_dist, int *knn_index, const knnhip_options &o, size_t budget,
                  bool kth_only);

// Learned router inference (knn_router_host.hip): the first probe of each
// query, into top1 (nullable) and/or bound (top1, q) = -inf of the tiled bound
// matrix lb (lb_at; nullable).
bool router_pick_first(knnhip_router *r, const float *q, int Q, long long ld, const int *counts,
                       int *top1, float *lb, hipStream_t s);
bool router_scores(knnhip_router *r, const float *q, int Q, long long ld, float *scores, hipStream_t s);
int router_dim(const knnhip_router *r);
int router_nblocks(const knnhip_router *r);

// Whole-file readers and writers of the index and router file formats:
// `ok` turns false at the first short read or write and stays false.
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

#define KNN_TRY(expr, what)                                                        \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess) {                                                    \
            knnhip::set_error("%s: %s", what, hipGetErrorString(_e));              \
            return false;                                                          \
        }                                                                          \
    } while (0)

}  // namespace knnhip

// The opaque index handle of the C API.
struct knnhip_index {
    int R = 0, dim = 0, nb = 0, pm_ld = 0, oct = 0, dpad = 0, ks = 0, nchunks = 1;
    int bmode = knnhip::kBRegs;
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
    // per ball, rebuilt from fnorm / frad: sqrt of the centre norm and the
    // radius, both rounded up to float
    knnhip::DevBuf fsq, fradf;
    // bfloat16 hi/lo tiles of the block-ordered references for the precision
    // scan, built by the first search that asks for it; nch_bf chunks
    int nch_bf = 0;
    knnhip::DevBuf packed_bf;
    // bfloat16 hi/lo ball centres for the three-term bound products, built
    // by the first precision search; fcent_nch chunks
    int fcent_nch = 0;
    knnhip::DevBuf fcent_bf;
    // slot layout of the bounds (idx_ball_lb_slots): nsl slots (0: a block
    // holds more than 32 balls, the per-ball fold is used), the ball of every
    // slot, per tile its mode word, per group of four slots its block; the
    // per-slot centres, norms, roots and radii are gathered by the first
    // search (sl_ready), the bfloat16 centres for sl_nch chunks
    int nsl = 0;
    bool sl_ready = false;
    int sl_nch = 0;
    knnhip::DevBuf sl_ball, sl_mode, sl_blk, sl_cent, sl_nrm, sl_sq, sl_rad, sl_bf;
    // position range of every ball (block-sorted order): its members occupy
    // positions ball_pos[f] .. ball_pos[f] + ball_cnt[f]; the ranges of a
    // block's balls partition the block.  Known for built indexes and for
    // files that store them; without them searches probe whole blocks only.
    bool ball_ranges = false;
    int max_bpb = 0;  // most balls in one block
    knnhip::DevBuf ball_pos, ball_cnt;
    // residual form: per unit (a ball where the ranges are known, else a
    // block) its centroid (rows of pm_ld floats) and radius, and the bfloat16
    // residual tiles of every position over res_nch chunks with their squared
    // norms; built by the first search that asks
    int res_nch = 0;
    bool res_balls = false;
    knnhip::DevBuf res_cent, res_tiles, res_norm, res_rho;
    // last search statistics
    int last_probes = 0;
    int last_fail = 0;
    long long last_tasks = 0;
    // the last search's tile bytes per reference and products (0 FP32, 1
    // three-term bfloat16, 2 residual bfloat16), for knnhip_index_last_work
    long long last_ref_bytes = 0;
    int last_prec = 0;
    // the last search's scan units (balls or whole blocks) and heap rows per query
    bool last_balls = false;
    int last_rows = 0;
    float last_ms[6] = {-1, -1, -1, -1, -1, -1};
    float rank_ms[3] = {-1, -1, -1};  // query norms, block bounds, block choice
};
