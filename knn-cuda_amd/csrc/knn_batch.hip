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
// Batched search: many independent problems, one persistent scan launch for
// all of them.  Problems are grouped into (dim, k) classes; within a class
// the problems' reference sets are laid out one after another (each starting
// on a 32-point tile) in one packed copy, and its tasks are (one slice of one
// problem's references, up to 32 of that problem's queries).  The scan's
// work queue holds every task of every class, largest first, and each block
// runs a task with its class's arguments and scan instance.  The rows-mode
// re-rank (one launch per class) maps positions back to each problem's own
// indexes, so every query gets exactly the result of a search of its own
// problem.  The slice length (references per task) is the class's tile
// shape, picked by the autotuner or, untuned, by the cost model below.
// ---------------------------------------------------------------------------
struct BatchWork {
    DevBuf packed, pm, norms, perm, rmax, starts, counts, query, qpm, qnorm, theta, tasks, qlist,
        rows, pval, pidx, dist, idx, fb_count, fb_list, qsub, dsub, isub, flist, rjobs, qjobs, rwg, qwg,
        outs;
    bool need(DevBuf &b, size_t bytes) { return b.n >= bytes || b.alloc(bytes); }
};
constexpr int kBatchSlices[] = {2048, 8192, 32768, 131072, 1 << 20};
// Largest task of a batched launch as a share of the work per resident
// block: untuned, and the shares the autotuner times.
constexpr double kBatchFrac = 0.5;
constexpr double kBatchTuneFracs[] = {0.125, 0.25, 0.5, 1.0, 2.0};

struct BatchTuneKey {
    int dim, k;
    bool operator<(const BatchTuneKey &o) const { return dim != o.dim ? dim < o.dim : k < o.k; }
};
std::map<BatchTuneKey, int> g_batch_table;  // (dim, k) -> slice length

struct BatchClass {
    int dim = 0, k = 0;
    std::vector<const knnhip_problem *> probs;
};

// Workspaces of a batched run: one per class, and the run's work queue.
std::vector<BatchWork> g_bwc;

// Pinned host memory for the run's uploads (task lists, query lists, rows,
// job tables): each array is copied into it and sent from there, a DMA
// transfer, where a copy from pageable memory goes through a staging kernel.
// Blocks are kept across runs; a run starts after the previous one's copies
// are done.
struct PinnedArena {
    std::vector<std::pair<unsigned char *, size_t>> blocks;
    size_t bi = 0, used = 0;
    void reset() {
        bi = 0;
        used = 0;
    }
    unsigned char *take(size_t n) {
        n = (n + 255) & ~(size_t)255;
        while (bi < blocks.size() && used + n > blocks[bi].second) {
            ++bi;
            used = 0;
        }
        if (bi == blocks.size()) {
            const size_t sz = std::max(n, (size_t)16 << 20);
            void *p = nullptr;
            if (hipHostMalloc(&p, sz, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                return nullptr;
            }
            blocks.emplace_back((unsigned char *)p, sz);
            used = 0;
        }
        unsigned char *p = blocks[bi].first + used;
        used += n;
        return p;
    }
};
PinnedArena g_pinned;

// dst (device) <- bytes of src (host), through the pinned arena, on stream s.
bool stage_copy(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return true;
    unsigned char *h = g_pinned.take(bytes);
    if (!h) {
        set_error("pinned allocation failed");
        return false;
    }
    memcpy(h, src, bytes);
    KNN_TRY(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, s), "copy");
    return true;
}
struct MultiWork {
    DevBuf cls, variant, work, next, nfail;
    int *nfail_host = nullptr;  // pinned, one per class
    int nfail_cap = 0;
    bool need(DevBuf &b, size_t bytes) { return b.n >= bytes || b.alloc(bytes); }
};
MultiWork g_mw;

// One class of a batched run.
struct ClassRun {
    const BatchClass *c = nullptr;
    bool direct = false;  // searched problem by problem, already finished
    Plan pl;
    int variant = 0;
    int L = 0, P = 0, ntasks = 0;
    long long Qtot = 0;
    std::vector<int> qstart;
    std::vector<ScanTask> tasks;
    ScanArgs sa{};
};

int batch_max_slices(const BatchClass &c, int L) {
    int m = 1;
    for (const knnhip_problem *p : c.probs) m = std::max(m, (p->ref_nb + L - 1) / L);
    return m;
}

// The batched scan's instance and padded dimension for a class's dim.
int batch_dpad(int dim) {
    int ks = 0, nch = 1;
    const int v = scan_multi_variant(dim, &ks, &nch);
    return v < 4 ? 2 * ks : 128 * nch;
}

// Resident blocks of the batched scan for these classes (its LDS is sized
// by their largest kp).
int batch_multi_resident(const std::vector<const BatchClass *> &cs, const knnhip_options &o) {
    size_t lds = scan_multi_lds(1);
    for (const BatchClass *c : cs) {
        const Plan pl = make_plan(1 << 20, kQueriesPerWave, c->dim, c->k, o);
        if (pl.use_scan) lds = std::max(lds, scan_multi_lds(pl.kp));
    }
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) {
        (void)hipGetLastError();
        cus = 256;
    }
    return cus * std::max(1, scan_multi_blocks_per_cu(lds));
}

// Slice lengths of one batched launch over every class: per class the
// longest slice whose largest task (128-reference steps plus ~3 steps of
// set-up, times the padded dimension) stays within `frac` of the work each
// resident block gets when all tasks of all classes are shared among them,
// within the re-rank's union (rows * kp <= kMaxUnion).  All tasks queue in
// one launch, so longer slices (fewer tasks, fewer heap rows to merge) only
// have to leave no single task longer than a block's share.
std::vector<int> batch_pick_slices(const std::vector<const BatchClass *> &cs, double frac, int resident,
                                   const knnhip_options &o) {
    const int Lmax = kBatchSlices[sizeof(kBatchSlices) / sizeof(int) - 1];
    auto task_cost = [](int span, int dpad) {
        return ((long long)(span + kRefsPerStep - 1) / kRefsPerStep + 3) * dpad;
    };
    double total = 0.0;
    for (const BatchClass *c : cs) {
        const int dpad = batch_dpad(c->dim);
        for (const knnhip_problem *p : c->probs)
            total += (double)((p->query_nb + kQueriesPerWave - 1) / kQueriesPerWave) *
                     task_cost(std::min(p->ref_nb, Lmax), dpad);
    }
    const double budget = frac * total / std::max(1, resident);
    std::vector<int> out;
    for (const BatchClass *c : cs) {
        const Plan pl = make_plan(1 << 20, kQueriesPerWave, c->dim, c->k, o);
        const int kp = std::max(pl.kp, c->k), dpad = batch_dpad(c->dim);
        int maxR = 0;
        for (const knnhip_problem *p : c->probs) maxR = std::max(maxR, p->ref_nb);
        int pick = 0;
        for (int L : kBatchSlices) {
            if ((long long)batch_max_slices(*c, L) * kp > kMaxUnion) continue;
            if (!pick) pick = L;  // the shortest the union allows
            if ((double)task_cost(std::min(L, maxR), dpad) <= budget) pick = L;
        }
        out.push_back(pick ? pick : Lmax);
    }
    return out;
}

// A class whose k leaves no room for the scan's heaps (searched problem by
// problem through the plain engine).
bool batch_class_direct(const BatchClass &c, const knnhip_options &o) {
    int maxR = 0;
    for (const knnhip_problem *p : c.probs) maxR = std::max(maxR, p->ref_nb);
    return !make_plan(maxR, kQueriesPerWave, c.dim, c.k, o).use_scan;
}

// Class c's layouts on stream s in workspace w, with the scan arguments and
// tasks for slice length L kept in run (the scan itself is the run's one
// batched launch).  A class whose k leaves no room for the scan's heaps is
// searched problem by problem at once (run.direct).
bool batch_class_prepare(const BatchClass &c, int L, const knnhip_options &o, BatchWork &w, hipStream_t s,
                         ClassRun &run) {
    run = ClassRun{};
    run.c = &c;
    const int dim = c.dim, k = c.k;
    int maxR = 0;
    for (const knnhip_problem *p : c.probs) maxR = std::max(maxR, p->ref_nb);
    Plan pl = make_plan(maxR, kQueriesPerWave, dim, k, o);
    if (!pl.use_scan) {
        for (const knnhip_problem *p : c.probs)
            if (!run_device(p->ref, p->ref_ld, p->ref_nb, p->query, p->query_ld, p->query_nb, dim, k,
                            p->knn_dist, p->knn_index, make_plan(p->ref_nb, p->query_nb, dim, k, o), o, s))
                return false;
        run.direct = true;
        return true;
    }
    // the batched scan's instance for this dimension and its packed layout
    int ks = 0, nch = 1;
    run.variant = scan_multi_variant(dim, &ks, &nch);
    pl.ks = ks;
    pl.nchunks = nch;
    pl.dpad = run.variant < 4 ? 2 * ks : 128 * nch;
    pl.oct = pl.dpad / 8;
    run.pl = pl;
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
    std::vector<ScanTask> &tasks = run.tasks;
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
    if (!stage_copy(w.starts.p, starts.data(), (size_t)np * 4, s) ||
        !stage_copy(w.counts.p, counts.data(), (size_t)np * 4, s) ||
        !stage_copy(w.tasks.p, tasks.data(), (size_t)ntasks * sizeof(ScanTask), s) ||
        !stage_copy(w.qlist.p, qlist.data(), nrow * 4, s) || !stage_copy(w.rows.p, rows.data(), rows.size() * 4, s))
        return false;
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
    if (!stage_copy(w.rjobs.p, rjobs.data(), (size_t)np * sizeof(NormJob), s) ||
        !stage_copy(w.qjobs.p, qjobs.data(), (size_t)np * sizeof(NormJob), s) ||
        !stage_copy(w.rwg.p, rwg.data(), rwg.size() * sizeof(int2), s) ||
        !stage_copy(w.qwg.p, qwg.data(), qwg.size() * sizeof(int2), s))
        return false;
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
    // the class's dim-major query block (read by the chunked instance) from the rows
    KNN_TRY(idx_transpose(w.qpm.as<float>(), pm_ld, (int)Qtot, dim, w.query.as<float>(), s), "queries");
    KNN_TRY(hipMemsetD32Async((hipDeviceptr_t)w.theta.p, 0xFF800000u, Qtot, s), "memset");
    KNN_TRY(hipMemsetAsync(w.fb_count.p, 0, 4, s), "memset");
    ScanArgs &sa = run.sa;
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
    run.L = L;
    run.P = P;
    run.ntasks = ntasks;
    run.Qtot = Qtot;
    run.qstart = qstart;
    return true;
}

// Class c's re-rank and certificate after the batched scan; its failure
// count goes to *nfail_dev.
bool batch_class_select(const ClassRun &run, const knnhip_options &o, BatchWork &w, hipStream_t s) {
    const BatchClass &c = *run.c;
    FinalArgs fa{};
    fa.R = run.sa.R;
    fa.query = w.query.as<float>();
    fa.ldq = run.Qtot;
    fa.Q = (int)run.Qtot;
    fa.dim = c.dim;
    fa.k = c.k;
    fa.kp = run.pl.kp;
    fa.nrows = run.P;
    fa.rows = w.rows.as<int>();
    fa.perm = w.perm.as<int>();
    fa.maxc = std::min(kMaxCand, std::max(2 * run.pl.kp, 256));
    fa.part_val = w.pval.as<float>();
    fa.part_idx = w.pidx.as<int>();
    fa.ref_pm = w.pm.as<float>();
    fa.query_pm = w.qpm.as<float>();
    fa.pm_ld = (c.dim + 3) & ~3;
    fa.qnorm = w.qnorm.as<float>();
    fa.rmax = w.rmax.as<unsigned>();
    fa.theta = w.theta.as<unsigned>();
    fa.force_fallback = o.force_fallback;
    fa.out_dist = w.dist.as<float>();
    fa.out_idx = w.idx.as<int>();
    fa.fb_count = w.fb_count.as<int>();
    fa.fb_list = w.fb_list.as<int>();
    KNN_TRY(launch_select(fa, s), "select kernel launch");
    return true;
}

// Finish: certificate failures (nfail of them) searched again (the plain
// search of their own problem), then each problem's k x query_nb outputs.
bool batch_class_finish(const ClassRun &run, int nfail, const knnhip_options &o, BatchWork &w, hipStream_t s) {
    const BatchClass &c = *run.c;
    const int dim = c.dim, k = c.k, np = (int)c.probs.size();
    const long long Qtot = run.Qtot;
    const std::vector<int> &qstart = run.qstart;
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
    if (!stage_copy(w.outs.p, outs.data(), (size_t)np * sizeof(BatchOut), s)) return false;
    KNN_TRY(idx_batch_outputs(w.dist.as<float>(), w.idx.as<int>(), k, (int)Qtot, w.outs.as<BatchOut>(), np, s),
            "outputs");
    return true;
}

// Streams of a batched run: the classes' layouts, re-ranks and outputs are
// spread over kBatchLanes streams (small launches of different classes
// overlap); the scan is one launch on the caller's stream between them.
constexpr int kBatchLanes = 4;
struct BatchLanes {
    bool ready = false;
    hipStream_t stream[kBatchLanes] = {};
    hipEvent_t done[kBatchLanes] = {};
    hipEvent_t start = nullptr;
};
BatchLanes g_lanes;

bool batch_lanes_ready() {
    if (g_lanes.ready) return true;
    for (int l = 0; l < kBatchLanes; ++l) {
        KNN_TRY(hipStreamCreateWithFlags(&g_lanes.stream[l], hipStreamNonBlocking), "stream create");
        KNN_TRY(hipEventCreateWithFlags(&g_lanes.done[l], hipEventDisableTiming), "event create");
    }
    KNN_TRY(hipEventCreateWithFlags(&g_lanes.start, hipEventDisableTiming), "event create");
    g_lanes.ready = true;
    return true;
}

// The lanes wait for what is queued on s so far.
bool batch_lanes_follow(hipStream_t s) {
    KNN_TRY(hipEventRecord(g_lanes.start, s), "event");
    for (int l = 0; l < kBatchLanes; ++l) KNN_TRY(hipStreamWaitEvent(g_lanes.stream[l], g_lanes.start, 0), "event wait");
    return true;
}

// s waits for what is queued on the lanes so far.
bool batch_lanes_join(hipStream_t s) {
    for (int l = 0; l < kBatchLanes; ++l) {
        KNN_TRY(hipEventRecord(g_lanes.done[l], g_lanes.stream[l]), "event");
        KNN_TRY(hipStreamWaitEvent(s, g_lanes.done[l], 0), "event wait");
    }
    return true;
}

// One batched run over classes cs with slice lengths Ls on stream s: every
// class's layouts, one persistent scan launch over all their tasks (timed
// into *scan_ms when asked), per-class re-rank, then failures and outputs.
// Classes searched problem by problem (k too large for the heaps) go first,
// on s alone: they use the plain engine's shared workspace, as do the
// re-searches of failed queries, which run one class at a time.
bool batch_run(const std::vector<const BatchClass *> &cs, const std::vector<int> &Ls, const knnhip_options &o,
               hipStream_t s, float *scan_ms) {
    const int n = (int)cs.size();
    if ((int)g_bwc.size() < n) g_bwc.resize(n);
    if (!batch_lanes_ready()) return false;
    // the previous run's uploads have left the pinned arena
    for (int l = 0; l < kBatchLanes; ++l) KNN_TRY(hipStreamSynchronize(g_lanes.stream[l]), "sync");
    KNN_TRY(hipStreamSynchronize(s), "sync");
    g_pinned.reset();
    std::vector<ClassRun> runs(n);
    std::vector<int> lane_of(n, 0);
    int nl = 0;
    for (int i = 0; i < n; ++i)
        if (batch_class_direct(*cs[i], o) && !batch_class_prepare(*cs[i], Ls[i], o, g_bwc[i], s, runs[i]))
            return false;
    if (!batch_lanes_follow(s)) return false;
    for (int i = 0; i < n; ++i) {
        if (runs[i].direct) continue;
        lane_of[i] = nl++ % kBatchLanes;
        if (!batch_class_prepare(*cs[i], Ls[i], o, g_bwc[i], g_lanes.stream[lane_of[i]], runs[i])) return false;
    }
    if (!batch_lanes_join(s)) return false;
    // the work queue: every task of every scanned class, larger first
    std::vector<ScanArgs> args;
    std::vector<int> variants, slot(n, -1);
    std::vector<std::pair<long long, int2>> items;
    size_t lds = 0;
    for (int i = 0; i < n; ++i) {
        const ClassRun &r = runs[i];
        if (r.direct || r.ntasks == 0) continue;
        slot[i] = (int)args.size();
        args.push_back(r.sa);
        variants.push_back(r.variant);
        lds = std::max(lds, scan_multi_lds(r.pl.kp));
        for (int t = 0; t < r.ntasks; ++t) {
            const ScanTask &tk = r.tasks[t];
            const long long steps = (tk.ref_end - tk.ref_begin + kRefsPerStep - 1) / kRefsPerStep;
            items.push_back({(steps + 3) * r.pl.dpad, make_int2(slot[i], t)});
        }
    }
    std::stable_sort(items.begin(), items.end(),
                     [](const std::pair<long long, int2> &x, const std::pair<long long, int2> &y) {
                         return x.first > y.first;
                     });
    const int nwork = (int)items.size();
    const int ncls = (int)args.size();
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (nwork > 0) {
        std::vector<int2> work(nwork);
        for (int i = 0; i < nwork; ++i) work[i] = items[i].second;
        if (!g_mw.need(g_mw.cls, (size_t)ncls * sizeof(ScanArgs)) || !g_mw.need(g_mw.variant, (size_t)ncls * 4) ||
            !g_mw.need(g_mw.work, (size_t)nwork * sizeof(int2)) || !g_mw.need(g_mw.next, 4)) {
            set_error("Memory allocation error");
            return false;
        }
        if (!stage_copy(g_mw.cls.p, args.data(), (size_t)ncls * sizeof(ScanArgs), s) ||
            !stage_copy(g_mw.variant.p, variants.data(), (size_t)ncls * 4, s) ||
            !stage_copy(g_mw.work.p, work.data(), (size_t)nwork * sizeof(int2), s))
            return false;
        KNN_TRY(hipMemsetAsync(g_mw.next.p, 0, 4, s), "memset");
        const int per_cu = scan_multi_blocks_per_cu(lds);
        if (per_cu < 1) {
            set_error("batched scan does not fit a CU (LDS %zu bytes)", lds);
            return false;
        }
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) {
            (void)hipGetLastError();
            cus = 256;
        }
        const int grid = std::min(nwork, cus * per_cu);
        if (scan_ms) {
            KNN_TRY(hipEventCreate(&e0), "event");
            KNN_TRY(hipEventCreate(&e1), "event");
            KNN_TRY(hipEventRecord(e0, s), "event");
        }
        KNN_TRY(launch_scan_multi(g_mw.cls.as<ScanArgs>(), g_mw.variant.as<int>(), g_mw.work.as<int2>(), nwork,
                                  g_mw.next.as<unsigned>(), lds, grid, s),
                "batched scan launch");
        if (scan_ms) KNN_TRY(hipEventRecord(e1, s), "event");
    }
    // re-rank per class on the lanes; the failure counts come back together
    if (g_mw.nfail_cap < n) {
        if (g_mw.nfail_host) (void)hipHostFree(g_mw.nfail_host);
        g_mw.nfail_host = nullptr;
        g_mw.nfail_cap = 0;
        KNN_TRY(hipHostMalloc((void **)&g_mw.nfail_host, (size_t)n * sizeof(int), hipHostMallocDefault),
                "pinned allocation");
        g_mw.nfail_cap = n;
    }
    if (!batch_lanes_follow(s)) return false;
    for (int i = 0; i < n; ++i) {
        g_mw.nfail_host[i] = 0;
        if (runs[i].direct || runs[i].ntasks == 0) continue;
        const hipStream_t ls = g_lanes.stream[lane_of[i]];
        if (!batch_class_select(runs[i], o, g_bwc[i], ls)) return false;
        KNN_TRY(hipMemcpyAsync(g_mw.nfail_host + i, g_bwc[i].fb_count.p, 4, hipMemcpyDeviceToHost, ls), "copy");
    }
    for (int l = 0; l < kBatchLanes; ++l) KNN_TRY(hipStreamSynchronize(g_lanes.stream[l]), "sync");
    if (scan_ms) {
        *scan_ms = 0.f;
        if (e0) {
            KNN_TRY(hipEventSynchronize(e1), "sync");
            KNN_TRY(hipEventElapsedTime(scan_ms, e0, e1), "event");
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
    }
    for (int i = 0; i < n; ++i) {
        if (runs[i].direct || runs[i].ntasks == 0) continue;
        if (!batch_class_finish(runs[i], g_mw.nfail_host[i], o, g_bwc[i], g_lanes.stream[lane_of[i]])) return false;
    }
    return batch_lanes_join(s);
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
    g_batch_fail = 0;
    std::string plan = "{\"classes\": [";
    bool first = true;
    const std::vector<BatchClass> classes = batch_classes(problems, count);
    std::vector<const BatchClass *> cs;
    std::vector<int> Ls;
    for (const BatchClass &c : classes) cs.push_back(&c);
    const std::vector<int> picked = batch_pick_slices(cs, kBatchFrac, batch_multi_resident(cs, o), o);
    for (size_t ci = 0; ci < classes.size(); ++ci) {
        const BatchClass &c = classes[ci];
        auto it = g_batch_table.find(BatchTuneKey{c.dim, c.k});
        const int L = it != g_batch_table.end() ? it->second : picked[ci];
        int ks = 0, nch = 1;
        const int v = scan_multi_variant(c.dim, &ks, &nch);
        char buf[192];
        snprintf(buf, sizeof(buf),
                 "%s{\"dim\": %d, \"k\": %d, \"problems\": %d, \"slice\": %d, \"tuned\": %s, \"instance\": %d}",
                 first ? "" : ", ", c.dim, c.k, (int)c.probs.size(), L, it != g_batch_table.end() ? "true" : "false",
                 v);
        plan += buf;
        first = false;
        Ls.push_back(L);
    }
    if (!batch_run(cs, Ls, o, s, nullptr)) return 0;
    KNN_TRY(hipStreamSynchronize(s), "sync");
    g_last_plan = plan + "], \"scan_launches\": 1}";
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
    // the whole batch in its one launch, for each share of a block's work a
    // task may take; the fastest sets every class's slice
    const std::vector<BatchClass> classes = batch_classes(problems, count);
    std::vector<const BatchClass *> cs;
    for (const BatchClass &c : classes) cs.push_back(&c);
    const int resident = batch_multi_resident(cs, o);
    std::vector<int> best;
    float best_ms = 0.f;
    for (double frac : kBatchTuneFracs) {
        const std::vector<int> Ls = batch_pick_slices(cs, frac, resident, o);
        if (!best.empty() && Ls == best) continue;
        float tot = 0.f;
        for (int r = 0; r < std::max(1, reps); ++r) {
            float ms = 0.f;
            if (!batch_run(cs, Ls, o, s, &ms)) return 0;
            tot += ms;
        }
        if (best.empty() || tot < best_ms) {
            best = Ls;
            best_ms = tot;
        }
    }
    for (size_t i = 0; i < classes.size(); ++i) g_batch_table[BatchTuneKey{classes[i].dim, classes[i].k}] = best[i];
    KNN_TRY(hipStreamSynchronize(s), "sync");
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
