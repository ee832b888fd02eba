// Host side of the batched search and its slice-length autotuner.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

#include "knn_engine.h"
#include "knn_index.h"

namespace knnhip {
namespace {

long long g_batch_fail = 0;  // certificate failures of the last batched search

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
        rows, pval, pidx, dist, idx, fb_count, fb_list, qsub, dsub, isub, flist, rjobs, qjobs, rwg, qwg,
        outs;
    bool need(DevBuf &b, size_t bytes) { return b.n >= bytes || b.alloc(bytes); }
};
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

// Classes in flight: each on its own stream and workspace.  A class is
// queued (layouts, scan, select, an asynchronous read-back of its failure
// count) and finished (failures searched again, outputs scattered) only when
// its lane comes round again, so the host prepares the next classes while
// the GPU runs the earlier ones, and small classes share the chip.
constexpr int kBatchLanes = 4;
BatchWork g_bw[kBatchLanes];
struct BatchLanes {
    bool ready = false;
    hipStream_t stream[kBatchLanes] = {};
    hipEvent_t counted[kBatchLanes] = {};  // failure count copied to host
    hipEvent_t start = nullptr;            // the caller's inputs are ready
    int *nfail = nullptr;                  // pinned, one per lane
};
BatchLanes g_lanes;

bool batch_lanes_ready() {
    if (g_lanes.ready) return true;
    for (int l = 0; l < kBatchLanes; ++l) {
        KNN_TRY(hipStreamCreateWithFlags(&g_lanes.stream[l], hipStreamNonBlocking), "stream create");
        KNN_TRY(hipEventCreateWithFlags(&g_lanes.counted[l], hipEventDisableTiming), "event create");
    }
    KNN_TRY(hipEventCreateWithFlags(&g_lanes.start, hipEventDisableTiming), "event create");
    KNN_TRY(hipHostMalloc((void **)&g_lanes.nfail, kBatchLanes * sizeof(int), hipHostMallocDefault),
            "pinned allocation");
    g_lanes.ready = true;
    return true;
}

// Lane streams wait for what the caller queued on its stream.
bool batch_lanes_follow(hipStream_t s) {
    KNN_TRY(hipEventRecord(g_lanes.start, s), "event");
    for (int l = 0; l < kBatchLanes; ++l) KNN_TRY(hipStreamWaitEvent(g_lanes.stream[l], g_lanes.start, 0), "event wait");
    return true;
}

// A class between its two phases.
struct ClassRun {
    const BatchClass *c = nullptr;
    int lane = 0;
    bool active = false;
    bool direct = false;  // searched problem by problem, already finished
    long long Qtot = 0;
    std::vector<int> qstart;
    hipEvent_t e0 = nullptr, e1 = nullptr;  // scan timing, when asked for
    float *scan_ms = nullptr;
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
int batch_default_slice(const BatchClass &c, int kp, int resident, int waves) {
    int best = 0;
    double best_cost = 1e300;
    for (int L : kBatchSlices) {
        if ((long long)batch_max_slices(c, L) * kp > kMaxUnion) continue;
        double cost = 0.0;
        // per-task work: a task's waves share its slice's 128-ref steps
        long long tasks = 0, steps = 0;
        for (const knnhip_problem *p : c.probs) {
            const long long ns = (p->ref_nb + L - 1) / L, ng = (p->query_nb + 31) / 32;
            tasks += ns * ng;
            const long long span = std::min<long long>(L, p->ref_nb);
            steps += ns * ng * ((span + kRefsPerStep - 1) / kRefsPerStep);
        }
        const double avg = (double)steps / std::max<long long>(1, tasks) / waves + 3.0;
        cost = std::ceil((double)tasks / resident) * avg;
        if (cost < best_cost) {
            best_cost = cost;
            best = L;
        }
    }
    return best;
}

// Queue phase of class c on `lane` (its stream and workspace; the lane's
// previous class must be finished).
bool batch_class_queue(const BatchClass &c, int L, const knnhip_options &o, int lane, float *scan_ms,
                       ClassRun &run) {
    BatchWork &w = g_bw[lane];
    const hipStream_t s = g_lanes.stream[lane];
    // the lane's previous outputs are out before its buffers are reused
    KNN_TRY(hipStreamSynchronize(s), "sync");
    run = ClassRun{};
    run.c = &c;
    run.lane = lane;
    run.active = true;
    run.scan_ms = scan_ms;
    const int dim = c.dim, k = c.k;
    int maxR = 0;
    for (const knnhip_problem *p : c.probs) maxR = std::max(maxR, p->ref_nb);
    const Plan pl = make_plan(maxR, kQueriesPerWave, dim, k, o);
    if (!pl.use_scan) {
        // k too large for the scan's heaps: each problem on its own, through
        // the plain engine's shared workspace, so synchronously
        for (const knnhip_problem *p : c.probs)
            if (!run_device(p->ref, p->ref_ld, p->ref_nb, p->query, p->query_ld, p->query_nb, dim, k,
                            p->knn_dist, p->knn_index, make_plan(p->ref_nb, p->query_nb, dim, k, o), o,
                            s))
                return false;
        KNN_TRY(hipStreamSynchronize(s), "sync");
        run.direct = true;
        if (scan_ms) *scan_ms = 0.f;
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
                tk.ref_lo = tk.ref_begin;
                tk.unit = i;
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
    // references and queries: one norms launch each over all the class's
    // sets (norms, point-major rows, and the references' packed tiles at
    // each set's start)
    std::vector<NormJob> rjobs(np), qjobs(np);
    std::vector<int2> rwg, qwg;
    for (int i = 0; i < np; ++i) {
        const knnhip_problem *p = c.probs[i];
        rjobs[i] = NormJob{p->ref, p->ref_ld, p->ref_nb,
                           (p->ref_nb + kRefsPerTile - 1) / kRefsPerTile * kRefsPerTile, starts[i]};
        qjobs[i] = NormJob{p->query, p->query_ld, p->query_nb, p->query_nb, qstart[i]};
        for (int i0 = 0; i0 < rjobs[i].n_pad; i0 += kNormPoints) rwg.push_back(make_int2(i, i0));
        for (int i0 = 0; i0 < p->query_nb; i0 += kNormPoints) qwg.push_back(make_int2(i, i0));
    }
    if (!w.need(w.rjobs, (size_t)np * sizeof(NormJob)) || !w.need(w.qjobs, (size_t)np * sizeof(NormJob)) ||
        !w.need(w.rwg, rwg.size() * sizeof(int2)) || !w.need(w.qwg, qwg.size() * sizeof(int2))) {
        set_error("Memory allocation error");
        return false;
    }
    KNN_TRY(hipMemcpyAsync(w.rjobs.p, rjobs.data(), (size_t)np * sizeof(NormJob), hipMemcpyHostToDevice, s),
            "copy");
    KNN_TRY(hipMemcpyAsync(w.qjobs.p, qjobs.data(), (size_t)np * sizeof(NormJob), hipMemcpyHostToDevice, s),
            "copy");
    KNN_TRY(hipMemcpyAsync(w.rwg.p, rwg.data(), rwg.size() * sizeof(int2), hipMemcpyHostToDevice, s), "copy");
    KNN_TRY(hipMemcpyAsync(w.qwg.p, qwg.data(), qwg.size() * sizeof(int2), hipMemcpyHostToDevice, s), "copy");
    NormArgs na{};
    na.dim = dim;
    na.oct = pl.oct;
    na.out = w.norms.as<float>();
    na.maxbits = w.rmax.as<unsigned>();
    na.pm = w.pm.as<float>();
    na.pm_ld = pm_ld;
    na.packed = w.packed.as<float>();
    KNN_TRY(launch_norms_jobs(na, w.rjobs.as<NormJob>(), w.rwg.as<int2>(), (int)rwg.size(), s),
            "norm kernel launch");
    KNN_TRY(idx_batch_perm(w.starts.as<int>(), w.counts.as<int>(), np, (int)n_pos, w.perm.as<int>(), s),
            "positions");
    NormArgs nq{};
    nq.dim = dim;
    nq.oct = (dim + 7) / 8;
    nq.out = w.qnorm.as<float>();
    nq.pm = w.qpm.as<float>();
    nq.pm_ld = pm_ld;
    KNN_TRY(launch_norms_jobs(nq, w.qjobs.as<NormJob>(), w.qwg.as<int2>(), (int)qwg.size(), s),
            "norm kernel launch");
    // the class's dim-major query block (read by the chunked scan) from the rows
    KNN_TRY(idx_transpose(w.qpm.as<float>(), pm_ld, (int)Qtot, dim, w.query.as<float>(), s), "queries");
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
    KNN_TRY(launch_scan(sa, pl.ks, pl.bmode, pl.waves, s), "scan kernel launch");
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
    // the failure count comes back asynchronously
    KNN_TRY(hipMemcpyAsync(g_lanes.nfail + lane, w.fb_count.p, 4, hipMemcpyDeviceToHost, s), "copy");
    KNN_TRY(hipEventRecord(g_lanes.counted[lane], s), "event");
    run.Qtot = Qtot;
    run.qstart = qstart;
    run.e0 = e0;
    run.e1 = e1;
    return true;
}

// Finish phase: certificate failures searched again (the plain search of
// their own problem), then each problem's k x query_nb outputs.
bool batch_class_finish(ClassRun &run, const knnhip_options &o) {
    if (!run.active) return true;
    run.active = false;
    if (run.direct) return true;
    const BatchClass &c = *run.c;
    const int lane = run.lane;
    BatchWork &w = g_bw[lane];
    const hipStream_t s = g_lanes.stream[lane];
    const int dim = c.dim, k = c.k, np = (int)c.probs.size();
    const long long Qtot = run.Qtot;
    const std::vector<int> &qstart = run.qstart;
    KNN_TRY(hipEventSynchronize(g_lanes.counted[lane]), "sync");
    const int nfail = g_lanes.nfail[lane];
    if (run.scan_ms) {
        KNN_TRY(hipEventSynchronize(run.e1), "sync");
        KNN_TRY(hipEventElapsedTime(run.scan_ms, run.e0, run.e1), "event");
        (void)hipEventDestroy(run.e0);
        (void)hipEventDestroy(run.e1);
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
    // each problem's k x query_nb outputs, one launch for the class
    std::vector<BatchOut> outs(np);
    for (int i = 0; i < np; ++i)
        outs[i] = BatchOut{c.probs[i]->knn_dist, c.probs[i]->knn_index, c.probs[i]->query_nb, qstart[i]};
    if (!w.need(w.outs, (size_t)np * sizeof(BatchOut))) {
        set_error("Memory allocation error");
        return false;
    }
    KNN_TRY(hipMemcpyAsync(w.outs.p, outs.data(), (size_t)np * sizeof(BatchOut), hipMemcpyHostToDevice, s),
            "copy");
    KNN_TRY(idx_batch_outputs(w.dist.as<float>(), w.idx.as<int>(), k, (int)Qtot, w.outs.as<BatchOut>(), np, s),
            "outputs");
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

// Resident scan blocks of a class's plan.
int batch_resident(const Plan &pl) { return std::max(1, pl.cus) * std::max(1, pl.blocks_per_cu); }

}  // namespace
}  // namespace knnhip

using namespace knnhip;

extern "C" {

int knnhip_search_batched(const knnhip_problem *problems, int count, const knnhip_options *opt,
                          void *stream) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!batch_args(problems, count) || !ensure_device()) return 0;
    const knnhip_options o = opt ? *opt : default_options(0);
    hipStream_t s = (hipStream_t)stream;
    CallOrder order(s);
    if (!batch_lanes_ready() || !batch_lanes_follow(s)) return 0;
    g_batch_fail = 0;
    std::string plan = "{\"classes\": [";
    bool first = true;
    const std::vector<BatchClass> classes = batch_classes(problems, count);
    ClassRun runs[kBatchLanes];
    int ci = 0;
    for (const BatchClass &c : classes) {
        const Plan pl = make_plan(1 << 20, kQueriesPerWave, c.dim, c.k, o);
        auto it = g_batch_table.find(BatchTuneKey{c.dim, c.k});
        int L = it != g_batch_table.end() ? it->second : batch_default_slice(c, std::max(pl.kp, c.k), batch_resident(pl), pl.waves);
        if (L == 0) L = kBatchSlices[sizeof(kBatchSlices) / sizeof(int) - 1];
        char buf[160];
        snprintf(buf, sizeof(buf), "%s{\"dim\": %d, \"k\": %d, \"problems\": %d, \"slice\": %d, \"tuned\": %s}",
                 first ? "" : ", ", c.dim, c.k, (int)c.probs.size(), L,
                 it != g_batch_table.end() ? "true" : "false");
        plan += buf;
        first = false;
        const int lane = ci++ % kBatchLanes;
        if (!batch_class_finish(runs[lane], o) || !batch_class_queue(c, L, o, lane, nullptr, runs[lane]))
            return 0;
    }
    for (int l = 0; l < kBatchLanes; ++l) {
        if (!batch_class_finish(runs[l], o)) return 0;
        if (hipStreamSynchronize(g_lanes.stream[l]) != hipSuccess) {
            set_error("sync: %s", hipGetErrorString(hipGetLastError()));
            return 0;
        }
    }
    g_last_plan = plan + "]}";
    g_last_fb = nullptr;
    g_last_nonfinite = nullptr;
    return 1;
}

long long knnhip_batched_last_failures(void) { return g_batch_fail; }

int knnhip_batched_autotune(const knnhip_problem *problems, int count, int reps, void *stream) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!batch_args(problems, count) || !ensure_device()) return 0;
    knnhip_options o = default_options(0);
    hipStream_t s = (hipStream_t)stream;
    CallOrder order(s);
    if (!batch_lanes_ready() || !batch_lanes_follow(s)) return 0;
    const std::vector<BatchClass> classes = batch_classes(problems, count);
    for (const BatchClass &c : classes) {
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
                ClassRun run;
                if (!batch_class_queue(c, L, o, 0, &ms, run) || !batch_class_finish(run, o)) return 0;
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
}  // extern "C"
