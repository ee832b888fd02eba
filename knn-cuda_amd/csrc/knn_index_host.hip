// Host side of the block index: build (k-means fine balls grouped into
// balanced blocks), search (ball lower bounds, block choice, task-mode scan,
// certificate with retry), and the index file format.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <utility>
#include <cmath>
#include <vector>

#include "knn_engine.h"
#include "knn_index.h"

namespace knnhip {
namespace {

constexpr int kIndexMargin = 8;  // block-index searches: fewer heap insertions per block
// ... with bfloat16 products, whose wider error bound needs a wider gap
// between the k-th and the kp-th candidate to certify
constexpr int kIndexMarginBf3 = 12;
// ... and with single-pass bfloat16 products on block residuals, whose bound
// grows with the block's radius
constexpr int kIndexMarginRes = 24;
// Waves per block-scan task.  Heap insertions of one query run one after
// another, so where the chip holds many tasks (kOneWaveQueries or more
// queries per block on average) one wave per task (up to eight tasks per CU,
// each with its own heaps) keeps more insertions in flight than four waves
// sharing one task's heaps; where tasks are few (a retry of a few queries)
// four waves per task shorten each task.
constexpr double kOneWaveQueries = 4.0;
// One-wave bfloat16 tasks per workgroup (each with its own heaps).
constexpr int kTasksPerGroup = 2;
// Ball probes: at most this many balls (heap rows) per query, the first
// included; further balls below the threshold are left to the certificate.
constexpr int kMaxBallRows = 24;

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

// Keeps the balls' position ranges (host arrays, block-sorted order) and the
// largest ball count of a block.
bool ball_ranges_set(knnhip_index *ix, const int *pos, const int *cnt, const int *fbs) {
    if (!ix->ball_pos.alloc((size_t)ix->nf * 4) || !ix->ball_cnt.alloc((size_t)ix->nf * 4)) {
        set_error("Memory allocation error");
        return false;
    }
    KNN_TRY(hipMemcpy(ix->ball_pos.p, pos, (size_t)ix->nf * 4, hipMemcpyHostToDevice), "copy");
    KNN_TRY(hipMemcpy(ix->ball_cnt.p, cnt, (size_t)ix->nf * 4, hipMemcpyHostToDevice), "copy");
    int m = 0;
    for (int b = 0; b < ix->nb; ++b) m = std::max(m, fbs[b + 1] - fbs[b]);
    ix->max_bpb = m;
    ix->ball_ranges = true;
    return true;
}

// Slot layout of the bounds from the blocks' ball ranges fbs (host, nb + 1):
// block b takes the smallest of 4, 8, 16, 32 slots that holds its balls,
// blocks by descending slot count (then by index), so each block starts at a
// multiple of its slot count; see ball_lb_slot_kernel.  A block of more than
// 32 balls leaves the index without a slot layout (nsl = 0).
bool ball_slots_set(knnhip_index *ix, const int *fbs) {
    ix->nsl = 0;
    ix->sl_ready = false;
    ix->sl_nch = 0;
    const int nb = ix->nb;
    std::vector<int> size(nb), order(nb);
    for (int b = 0; b < nb; ++b) {
        const int n = fbs[b + 1] - fbs[b];
        int sz = 4;
        while (sz < n) sz *= 2;
        if (sz > 32) return true;
        size[b] = sz;
        order[b] = b;
    }
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return size[x] > size[y]; });
    long long total = 0;
    for (int b = 0; b < nb; ++b) total += size[b];
    const int nsl = (int)((total + 31) / 32 * 32);
    const int ntile = nsl / 32;
    std::vector<int> ball(nsl, -1), blk((size_t)ntile * 8, -1), mode(ntile, 0);
    int pos = 0;
    for (int b : order) {
        const int sz = size[b];
        for (int f = fbs[b]; f < fbs[b + 1]; ++f) ball[pos + (f - fbs[b])] = f;
        const int t = pos / 32, o = pos % 32, j = o / 8, h = (o % 8) / 4;
        blk[(size_t)t * 8 + 2 * j + h] = b;
        const int code = sz == 4 ? 0 : sz == 8 ? 1 : sz == 16 ? 2 : 3;
        for (int g = j; g < j + std::max(1, sz / 8); ++g) mode[t] |= code << (2 * g);
        pos += sz;
    }
    auto upload = [](DevBuf &d, const std::vector<int> &v) {
        return d.alloc(v.size() * 4) && hipMemcpy(d.p, v.data(), v.size() * 4, hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!upload(ix->sl_ball, ball) || !upload(ix->sl_mode, mode) || !upload(ix->sl_blk, blk)) {
        set_error("Memory allocation error");
        return false;
    }
    ix->nsl = nsl;
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
        ix->bmode = p.bmode;
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
    if (!ix->fsq.alloc((size_t)nf * 4) || !ix->fradf.alloc((size_t)nf * 4)) {
        set_error("Memory allocation error");
        return false;
    }
    KNN_TRY(idx_ball_aux(ix->fnorm.as<float>(), ix->frad.as<double>(), nf, ix->fsq.as<float>(),
                         ix->fradf.as<float>(), s),
            "ball roots");
    // each ball's position range, in the block-sorted order of the balls
    {
        std::vector<int> hs(nf), hst(nf), hc(nf), hfb(nb + 1);
        KNN_TRY(hipMemcpyAsync(hs.data(), fsorted.p, (size_t)nf * 4, hipMemcpyDeviceToHost, s), "copy");
        KNN_TRY(hipMemcpyAsync(hst.data(), fstart.p, (size_t)nf * 4, hipMemcpyDeviceToHost, s), "copy");
        KNN_TRY(hipMemcpyAsync(hc.data(), fcount.p, (size_t)nf * 4, hipMemcpyDeviceToHost, s), "copy");
        KNN_TRY(hipMemcpyAsync(hfb.data(), ix->fb_start.p, (size_t)(nb + 1) * 4, hipMemcpyDeviceToHost, s),
                "copy");
        KNN_TRY(hipStreamSynchronize(s), "sync");
        std::vector<int> bp(nf), bc(nf);
        for (int j = 0; j < nf; ++j) {
            bp[j] = hst[hs[j]];
            bc[j] = hc[hs[j]];
        }
        if (!ball_ranges_set(ix, bp.data(), bc.data(), hfb.data())) return false;
        if (!ball_slots_set(ix, hfb.data())) return false;
    }
    KNN_TRY(hipEventRecord(e1, s), "event");
    KNN_TRY(hipEventSynchronize(e1), "event");
    KNN_TRY(hipEventElapsedTime(&ix->build_ms, e0, e1), "event");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return true;
}

// On-disk format of a block index (little-endian): magic "KNNHIPIX", uint32
// version (1 or 2), int32 R, dim, nb, nf, n_pos, r_pad; then block_start
// [nb+1], block_count [nb], perm [r_pad] (int32), ball centres [nf][pm_ld],
// centre norms [nf] (float32), radii [nf] (float64), ball blocks [nf] and
// ball starts [nb+1] (int32); version 2 adds each ball's first position and
// member count [nf] each (int32).  The packed tiles, point-major copy and
// norms are rebuilt from the caller's reference set when the file is loaded.
constexpr char kIndexMagic[8] = {'K', 'N', 'N', 'H', 'I', 'P', 'I', 'X'};
constexpr unsigned kIndexVersion = 2;

bool index_save(const knnhip_index *ix, const char *path) {
    FILE *f = fopen(path, "wb");
    if (!f) {
        set_error("cannot open %s for writing", path);
        return false;
    }
    FileOut o{f};
    const int hdr[6] = {ix->R, ix->dim, ix->nb, ix->nf, ix->n_pos, ix->r_pad};
    const unsigned version = ix->ball_ranges ? 2u : 1u;
    o.put(kIndexMagic, 8);
    o.put(&version, 4);
    o.put(hdr, sizeof(hdr));
    o.put_dev(ix->block_start, (size_t)(ix->nb + 1) * 4);
    o.put_dev(ix->block_count, (size_t)ix->nb * 4);
    o.put_dev(ix->perm, (size_t)ix->r_pad * 4);
    o.put_dev(ix->fcent, (size_t)ix->nf * ix->pm_ld * 4);
    o.put_dev(ix->fnorm, (size_t)ix->nf * 4);
    o.put_dev(ix->frad, (size_t)ix->nf * 8);
    o.put_dev(ix->fblock, (size_t)ix->nf * 4);
    o.put_dev(ix->fb_start, (size_t)(ix->nb + 1) * 4);
    if (ix->ball_ranges) {
        o.put_dev(ix->ball_pos, (size_t)ix->nf * 4);
        o.put_dev(ix->ball_cnt, (size_t)ix->nf * 4);
    }
    const bool ok = o.ok && fclose(f) == 0;
    if (!ok) set_error("cannot write %s", path);
    return ok;
}

// Structural checks of a loaded index, or nullptr when it is consistent:
// blocks start on whole tiles, in order, and end by n_pos; the first count
// positions of each block hold distinct reference indexes in [0, R) and the
// rest -1, so that every reference appears exactly once; ball ranges
// partition [0, nf) in block order, and every non-empty block has a ball;
// centres are finite.
const char *index_layout_error(int R, int dim, int nb, int nf, int n_pos, int r_pad, int pm_ld, const int *bs,
                               const int *bc, const int *perm, const int *fbs, const int *fbl, const float *fc) {
    if (r_pad != (n_pos + kRefsPerStep - 1) / kRefsPerStep * kRefsPerStep) return "padded size";
    if (bs[0] != 0 || bs[nb] != n_pos) return "block starts";
    if (fbs[0] != 0 || fbs[nb] != nf) return "ball ranges";
    std::vector<char> seen(R, 0);
    long long found = 0;
    for (int b = 0; b < nb; ++b) {
        if (bs[b] % kRefsPerTile != 0 || bc[b] < 0 || (long long)bs[b] + bc[b] > bs[b + 1])
            return "block starts";
        if (fbs[b + 1] < fbs[b] || (bc[b] > 0 && fbs[b + 1] == fbs[b])) return "ball ranges";
        for (int f = fbs[b]; f < fbs[b + 1]; ++f)
            if (fbl[f] != b) return "ball blocks";
        for (int p = bs[b]; p < bs[b + 1]; ++p) {
            const int v = perm[p];
            if (p < bs[b] + bc[b]) {
                if (v < 0 || v >= R || seen[v]) return "permutation";
                seen[v] = 1;
                ++found;
            } else if (v != -1) {
                return "permutation";
            }
        }
    }
    for (int p = n_pos; p < r_pad; ++p)
        if (perm[p] != -1) return "permutation";
    if (found != R) return "permutation";
    for (int f = 0; f < nf; ++f)
        for (int d = 0; d < dim; ++d)
            if (!std::isfinite(fc[(size_t)f * pm_ld + d])) return "ball centres";
    return nullptr;
}

// Ball position ranges of a loaded index, or nullptr when consistent: the
// ranges of a block's balls (in any order) tile exactly the block's first
// count positions.
const char *ball_range_error(int nb, const int *bs, const int *bc, const int *fbs, const int *bp,
                             const int *bn) {
    std::vector<std::pair<long long, long long>> r;
    for (int b = 0; b < nb; ++b) {
        r.clear();
        for (int f = fbs[b]; f < fbs[b + 1]; ++f) {
            if (bn[f] < 0) return "ball position ranges";
            r.emplace_back(bp[f], (long long)bp[f] + bn[f]);
        }
        std::sort(r.begin(), r.end());
        long long at = bs[b];
        for (const auto &x : r) {
            if (x.first != at) return "ball position ranges";
            at = x.second;
        }
        if (at != (long long)bs[b] + bc[b]) return "ball position ranges";
    }
    return nullptr;
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
    if (!in.ok || memcmp(magic, kIndexMagic, 8) != 0 || version < 1 || version > kIndexVersion) {
        fclose(f);
        set_error("%s is not a block index file of version 1 to %u", path, kIndexVersion);
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
    ix->bmode = p.bmode;
    ix->nchunks = p.nchunks;
    ix->dpad = p.dpad;
    ix->oct = p.oct;
    // The arrays pass through host memory and are checked before any kernel
    // indexes with them.  The centres' norms and the radii in the file are
    // not trusted: both are recomputed below from the centres and from the
    // reference set the index is loaded over.
    const int nb = ix->nb, nf = ix->nf;
    std::vector<int> bs(nb + 1), bc(nb), perm(ix->r_pad), fbl(nf), fbs(nb + 1);
    std::vector<float> fc((size_t)nf * ix->pm_ld), fn(nf);
    std::vector<double> fr(nf);
    in.get(bs.data(), bs.size() * 4);
    in.get(bc.data(), bc.size() * 4);
    in.get(perm.data(), perm.size() * 4);
    in.get(fc.data(), fc.size() * 4);
    in.get(fn.data(), fn.size() * 4);
    in.get(fr.data(), fr.size() * 8);
    in.get(fbl.data(), fbl.size() * 4);
    in.get(fbs.data(), fbs.size() * 4);
    std::vector<int> bp, bn;
    if (version >= 2) {
        bp.resize(nf);
        bn.resize(nf);
        in.get(bp.data(), bp.size() * 4);
        in.get(bn.data(), bn.size() * 4);
    }
    char extra;
    const bool tail = fread(&extra, 1, 1, f) == 0;
    fclose(f);
    if (!in.ok || !tail) {
        set_error("%s is truncated or has trailing data", path);
        return false;
    }
    if (const char *why = index_layout_error(R, dim, nb, nf, ix->n_pos, ix->r_pad, ix->pm_ld, bs.data(),
                                             bc.data(), perm.data(), fbs.data(), fbl.data(), fc.data())) {
        set_error("%s: inconsistent index (%s)", path, why);
        return false;
    }
    if (version >= 2) {
        if (const char *why = ball_range_error(nb, bs.data(), bc.data(), fbs.data(), bp.data(), bn.data())) {
            set_error("%s: inconsistent index (%s)", path, why);
            return false;
        }
        if (!ball_ranges_set(ix, bp.data(), bn.data(), fbs.data())) return false;
    } else {
        ix->ball_ranges = false;
    }
    if (!ball_slots_set(ix, fbs.data())) return false;
    auto upload = [](DevBuf &b, const void *src, size_t n) {
        return b.alloc(n) && hipMemcpy(b.p, src, n, hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!upload(ix->block_start, bs.data(), bs.size() * 4) || !upload(ix->block_count, bc.data(), bc.size() * 4) ||
        !upload(ix->perm, perm.data(), perm.size() * 4) || !upload(ix->fcent, fc.data(), fc.size() * 4) ||
        !upload(ix->fblock, fbl.data(), fbl.size() * 4) || !upload(ix->fb_start, fbs.data(), fbs.size() * 4) ||
        !ix->fnorm.alloc((size_t)nf * 4) || !ix->frad.alloc((size_t)nf * 8)) {
        set_error("Memory allocation error");
        return false;
    }
    // rebuild the layouts from the references
    DevBuf pm0, norms0, d2bits, member;
    if (!pm0.alloc((size_t)R * ix->pm_ld * 4) || !norms0.alloc((size_t)R * 4) || !ix->rmax.alloc(4) ||
        !ix->fsq.alloc((size_t)ix->nf * 4) || !ix->fradf.alloc((size_t)ix->nf * 4) ||
        !ix->packed.alloc((size_t)ix->r_pad * ix->dpad * 4) ||
        !ix->pm.alloc((size_t)ix->r_pad * ix->pm_ld * 4) || !ix->norms.alloc((size_t)ix->r_pad * 4) ||
        !d2bits.alloc((size_t)nf * 8) || !member.alloc((size_t)nf * 4)) {
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
    // balls: norms of the centres, radii over the points they now hold
    KNN_TRY(idx_row_norms(ix->fcent.as<float>(), ix->pm_ld, nf, dim, ix->fnorm.as<float>(), s), "norms");
    if (ix->ball_ranges) {
        // each ball keeps its stored members: radius over its own range
        KNN_TRY(idx_ball_range_radius(ix->pm.as<float>(), ix->pm_ld, dim, ix->fcent.as<float>(),
                                      ix->ball_pos.as<int>(), ix->ball_cnt.as<int>(), nf, ix->frad.as<double>(), s),
                "ball radii");
    } else {
        KNN_TRY(idx_ball_refit(ix->pm.as<float>(), ix->pm_ld, dim, ix->block_start.as<int>(),
                               ix->block_count.as<int>(), ix->fb_start.as<int>(), nb, ix->fcent.as<float>(), nf,
                               d2bits.as<unsigned long long>(), member.as<int>(), ix->frad.as<double>(), s),
                "ball radii");
    }
    KNN_TRY(idx_ball_aux(ix->fnorm.as<float>(), ix->frad.as<double>(), ix->nf, ix->fsq.as<float>(),
                         ix->fradf.as<float>(), s),
            "ball roots");
    KNN_TRY(hipStreamSynchronize(s), "sync");
    return true;
}

// Workspace of index searches (grown on demand, kept between calls).
struct IndexWork {
    DevBuf qnorm, qpm, lb, chosen, lb_rest, pcount, qstart, cursor, qlist, rows, tasks,
        pval, pidx, theta, fb_count, fb_list, qsub, dsub, isub, scores, resrow, bchosen;
    bool need(DevBuf &b, size_t bytes) { return b.n >= bytes || b.alloc(bytes); }
};
IndexWork g_iw, g_iw_retry;

// Probes of the retry that queries failing their certificate get before the
// full scan (ranked by the lower bounds alone).
constexpr int kRetryProbes = 32;

// Error coefficient of the block bounds with bfloat16 three-term products
// (replaces 1.5 g of the FP32 products in ball_lower_bound).  With u = 2^-8,
// n = 48 nch product terms accumulated in FP32 (unit 2^-23, allowing for any
// rounding inside the MFMA) and the FP32 norms of q and c (unit 2^-23 as
// well):
//   |nq + nc - 2 dot - |q - c|^2| <= g_D (|q|^2 + |c|^2)
//                                   + 2 (3.01 u^2 + 1.02 g_n) |q| |c|
//                                 <= (g_D + (3.01 u^2 + 1.02 g_n) / 2) S,
// S = (|q| + |c|)^2, since 2 |q| |c| <= S / 2; widened by 5 % and rounded up.
float ball_error_coeff_bf3(int dim, int nch) {
    const double u2 = 1.0 / 65536.0;
    const double n = 48.0 * nch;
    const double gn = n * 0x1p-23 / (1.0 - n * 0x1p-23);
    const double gd = dim * 0x1p-23 / (1.0 - dim * 0x1p-23);
    const double c = 1.05 * (gd + 0.5 * (3.01 * u2 + 1.02 * gn));
    return (float)c * (1.0f + 0x1p-20f);
}

// The bfloat16 ball centres of the precision bounds (built once per index).
bool ball_bf3_ready(knnhip_index *ix, int nch, hipStream_t s) {
    if (ix->fcent_nch == nch) return true;
    ix->fcent_nch = 0;
    if (!ix->fcent_bf.alloc((size_t)ix->nf * nch * 32 * 2)) {
        set_error("Memory allocation error");
        return false;
    }
    KNN_TRY(idx_pack_ball_bf3(ix->fcent.as<float>(), ix->pm_ld, ix->nf, ix->dim, nch,
                              ix->fcent_bf.as<unsigned short>(), s),
            "bfloat16 ball centres");
    ix->fcent_nch = nch;
    return true;
}

// The per-slot arrays of the slot-form bounds (once per index), and the
// bfloat16 slot centres for nch chunks.
bool ball_slots_ready(knnhip_index *ix, int nch, hipStream_t s) {
    const int nsl = ix->nsl;
    if (!ix->sl_ready) {
        if (!ix->sl_cent.alloc((size_t)nsl * ix->pm_ld * 4) || !ix->sl_nrm.alloc((size_t)nsl * 4) ||
            !ix->sl_sq.alloc((size_t)nsl * 4) || !ix->sl_rad.alloc((size_t)nsl * 4)) {
            set_error("Memory allocation error");
            return false;
        }
        KNN_TRY(idx_ball_slot_gather(ix->sl_ball.as<int>(), nsl, ix->fcent.as<float>(), ix->pm_ld,
                                     ix->fnorm.as<float>(), ix->fsq.as<float>(), ix->fradf.as<float>(),
                                     ix->sl_cent.as<float>(), ix->sl_nrm.as<float>(), ix->sl_sq.as<float>(),
                                     ix->sl_rad.as<float>(), s),
                "bound slots");
        ix->sl_ready = true;
    }
    if (nch && ix->sl_nch != nch) {
        ix->sl_nch = 0;
        if (!ix->sl_bf.alloc((size_t)nsl * nch * 32 * 2)) {
            set_error("Memory allocation error");
            return false;
        }
        KNN_TRY(idx_pack_ball_bf3(ix->sl_cent.as<float>(), ix->pm_ld, nsl, ix->dim, nch,
                                  ix->sl_bf.as<unsigned short>(), s),
                "bfloat16 slot centres");
        ix->sl_nch = nch;
    }
    return true;
}

// Block lower bounds of the Q queries whose norms and point-major rows are
// in qnorm / qpm, into lb (nb x Q, block-major); bfloat16 three-term
// products with `precision` where the dimension allows (<= 128).  The slot
// form where the index has one and the dimension fits it, unless per_ball
// asks for the per-ball fold (the same bounds).
bool index_bounds(knnhip_index *ix, const float *qpm, const float *qnorm, int Q, bool precision, float *lb,
                  hipStream_t s, bool per_ball = false) {
    const int nch = precision && ix->dim <= 128 ? (ix->dim + 15) / 16 : 0;
    if (!per_ball && ix->nsl > 0 && idx_ball_lb_slots_fit(ix->dim, ix->pm_ld, nch)) {
        if (!ball_slots_ready(ix, nch, s)) return false;
        KNN_TRY(idx_ball_lb_slots(qpm, ix->pm_ld, Q, ix->dim, qnorm, ix->sl_cent.as<float>(),
                                  nch ? ix->sl_bf.as<unsigned short>() : nullptr, nch,
                                  nch ? ball_error_coeff_bf3(ix->dim, nch) : 0.0f, ix->nsl, ix->sl_nrm.as<float>(),
                                  ix->sl_sq.as<float>(), ix->sl_rad.as<float>(), ix->sl_blk.as<int>(),
                                  ix->sl_mode.as<int>(), ix->nb, lb, s),
                "block bounds");
        return true;
    }
    if (nch && !ball_bf3_ready(ix, nch, s)) return false;
    KNN_TRY(idx_ball_lb(qpm, ix->pm_ld, Q, ix->dim, qnorm, ix->fcent.as<float>(), ix->nf,
                        ix->fnorm.as<float>(), ix->fsq.as<float>(), ix->fradf.as<float>(),
                        ix->fblock.as<int>(), ix->fb_start.as<int>(), ix->nb, lb, s,
                        nch ? ix->fcent_bf.as<unsigned short>() : nullptr, nch,
                        nch ? ball_error_coeff_bf3(ix->dim, nch) : 0.0f),
            "block bounds");
    return true;
}

// depth 0: a caller's search; depth 1: the retry of its failed queries.
bool index_search(knnhip_index *ix, IndexWork &w, int depth, const float *query, int Q,
                  long long ldq, int k, int P, const float *router_scores, float *out_dist,
                  int *out_idx, const knnhip_options &o, hipStream_t s, knnhip_router *router = nullptr) {
    const int nb = ix->nb, dim = ix->dim, pm_ld = ix->pm_ld;
    // precision 1: bfloat16 three-term products for dimensions up to 128;
    // precision 2: single-pass bfloat16 products on block residuals for
    // dimensions up to 256, else the three-term products where they apply;
    // otherwise FP32 products
    const int nch_res = o.precision == 2 ? scan_res_nch(dim) : 0;
    const int nch_bf = (o.precision == 1 || (o.precision == 2 && !nch_res)) ? scan_bf3_task_nch(dim) : 0;
    const int margin = o.margin > 0     ? o.margin
                       : k >= kLargeK   ? kDefaultMarginLargeK
                       : nch_res        ? kIndexMarginRes
                       : nch_bf         ? kIndexMarginBf3
                                        : kIndexMargin;
    const int kp = std::max(k, std::min(k + margin, ix->R));
    // wide dimensions: query rows staged in LDS where they fit beside this
    // search's heaps, else the chunked instance over the same tiles
    int bmode = ix->bmode, ks = ix->ks;
    if (bmode == kBLds && scan_block_lds(kp, 4, kBLds, ix->oct) > kScanLdsMax) {
        bmode = kBChunks;
        ks = 64;
    }
    if (scan_block_lds(kp, 4, bmode, ix->oct) > kScanLdsMax || (long long)P * kp > kMaxUnion) {
        set_error("k or probe count too large for the block index (k=%d probes=%d)", k, P);
        return false;
    }
    if (o.probe_units < 0 || o.probe_units > 2) {
        set_error("bad probe_units option (%d)", o.probe_units);
        return false;
    }
    P = std::max(1, std::min(P, nb));
    // Ball probes (where the index knows its balls' position ranges), Pb heap
    // rows per query.  One pass (probe_units 0): the balls of the P blocks
    // whose bounds lie below the reach of the nearest balls holding kp
    // members, at most Pb.  Two passes (2): the nearest ball of the first
    // block, then up to `cap` more balls whose bounds lie below the threshold
    // that scan left.  Whole blocks otherwise (1): P rows.
    const bool balls =
        ix->ball_ranges && o.probe_units != 1 && (long long)P * ix->max_bpb <= idx_ball_cand_max();
    const bool one_pass = balls && o.probe_units == 0;
    const int rows_max = std::min({P * ix->max_bpb, kMaxUnion / kp, kMaxBallRows});
    const int cap = balls && !one_pass ? std::max(0, rows_max - 1) : 0;
    const int Pb = one_pass ? std::max(1, rows_max) : balls ? 1 + cap : P;
    const int nu = balls ? ix->nf : nb;
    const int *unit_start = balls ? ix->ball_pos.as<int>() : ix->block_start.as<int>();
    const int *unit_count = balls ? ix->ball_cnt.as<int>() : ix->block_count.as<int>();
    const long long npairs = (long long)Q * Pb;
    if (nch_bf && ix->nch_bf != nch_bf) {
        ix->nch_bf = 0;
        if (!ix->packed_bf.alloc((size_t)ix->r_pad / kRefsPerTile * nch_bf * 2 * kWave * 16)) {
            set_error("Memory allocation error");
            return false;
        }
        KNN_TRY(launch_pack_bf3_rows(ix->pm.as<float>(), pm_ld, ix->r_pad, nch_bf,
                                     ix->packed_bf.as<unsigned short>(), s),
                "bfloat16 tiles");
        ix->nch_bf = nch_bf;
    }
    if (nch_res && (ix->res_nch != nch_res || ix->res_balls != balls)) {
        // residual tiles against the centre of each scan unit (ball or block)
        ix->res_nch = 0;
        const size_t ntile = (size_t)ix->r_pad / kRefsPerTile;
        if (!ix->res_cent.alloc((size_t)nu * pm_ld * 4) || !ix->res_tiles.alloc(ntile * nch_res * kWave * 16) ||
            !ix->res_norm.alloc((size_t)ix->r_pad * 4) || !ix->res_rho.alloc((size_t)nu * 4)) {
            set_error("Memory allocation error");
            return false;
        }
        KNN_TRY(hipMemsetAsync(ix->res_tiles.p, 0, ix->res_tiles.n, s), "memset");
        KNN_TRY(hipMemsetAsync(ix->res_norm.p, 0, ix->res_norm.n, s), "memset");
        KNN_TRY(launch_resid_build(ix->pm.as<float>(), pm_ld, dim, nch_res, unit_start, unit_count, nu,
                                   ix->res_cent.as<float>(), ix->res_tiles.as<unsigned short>(),
                                   ix->res_norm.as<float>(), ix->res_rho.as<float>(), s),
                "residual tiles");
        ix->res_nch = nch_res;
        ix->res_balls = balls;
    }
    // Two scan phases: each query's first unit (its first-ranked block, or
    // that block's nearest ball), then its other units, which start from the
    // thresholds the first phase left.  One-pass ball probes: one phase.
    const int nphase = one_pass ? 1 : Pb > 1 ? 2 : 1;
    const int pr[3] = {0, one_pass ? Pb : 1, Pb};
    // one-pass ball tasks of the bfloat16 forms keep merged candidate lists
    // (tasks of up to kMergeQueries queries); other tasks, heaps (32 queries)
    const bool merge = one_pass && (nch_bf || nch_res) && kp <= kMergeMaxKp;
    const int qpt = merge ? kMergeQueries : kQueriesPerWave;
    int max_t[2] = {0, 0};
    for (int ph = 0; ph < nphase; ++ph)
        max_t[ph] = (int)(((long long)Q * (pr[ph + 1] - pr[ph]) + qpt - 1) / qpt) + nu;
    const int max_tasks = max_t[0] + max_t[1];
    if (!w.need(w.qnorm, (size_t)Q * 4) || !w.need(w.qpm, (size_t)Q * pm_ld * 4) ||
        !w.need(w.lb, (size_t)lb_elems(Q, nb) * 4) ||
        !w.need(w.chosen, (size_t)Q * P * 4) || !w.need(w.lb_rest, (size_t)Q * 4) ||
        (balls && !w.need(w.bchosen, (size_t)npairs * 4)) ||
        !w.need(w.pcount, (size_t)nu * 4) || !w.need(w.qstart, (size_t)(nu + 1) * 4) ||
        !w.need(w.cursor, (size_t)(nu + 1) * 4) ||
        !w.need(w.qlist, (size_t)max_tasks * qpt * 4) ||
        !w.need(w.rows, (size_t)npairs * 4) || !w.need(w.tasks, (size_t)max_tasks * sizeof(ScanTask)) ||
        !w.need(w.pval, (size_t)max_tasks * qpt * kp * 4) ||
        !w.need(w.pidx, (size_t)max_tasks * qpt * kp * 4) ||
        !w.need(w.theta, (size_t)Q * 4) || !w.need(w.fb_count, 4) || !w.need(w.fb_list, (size_t)Q * 4) ||
        (nch_res && !w.need(w.resrow, (size_t)max_tasks * qpt * 16))) {
        set_error("Memory allocation error");
        return false;
    }
    // stage timing: device-side timestamps only (no system-scope fence, so
    // recording one does not stall the stream)
    // (timing 2: the scan stage's two events only)
    hipEvent_t ev[6], evr[2], evk[2][2];
    const bool timing = o.timing == 1, scan_timing = o.timing != 0;
    if (timing)
        for (auto &e : evr) KNN_TRY(hipEventCreateWithFlags(&e, hipEventDisableSystemFence), "event");
    if (scan_timing)
        for (auto &p : evk)
            for (auto &e : p) KNN_TRY(hipEventCreateWithFlags(&e, hipEventDisableSystemFence), "event");
    if (timing)
        for (auto &e : ev) KNN_TRY(hipEventCreateWithFlags(&e, hipEventDisableSystemFence), "event");
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
    if (!index_bounds(ix, w.qpm.as<float>(), w.qnorm.as<float>(), Q, o.precision >= 1, w.lb.as<float>(), s))
        return false;
    // scores pick the first probe (or rank every probe, with score_all); a
    // router handle picks it with the fused inference kernel
    if (router && o.score_all) {
        if (w.scores.n < (size_t)nb * Q * 4 && !w.scores.alloc((size_t)nb * Q * 4)) {
            set_error("Memory allocation error");
            return false;
        }
        if (!knnhip::router_scores(router, query, Q, ldq, w.scores.as<float>(), s)) return false;
        router_scores = w.scores.as<float>();
    }
    const float *score = o.score_all ? router_scores : nullptr;
    if (router && !o.score_all) {
        if (!router_pick_first(router, query, Q, ldq, ix->block_count.as<int>(), nullptr, w.lb.as<float>(), s))
            return false;
    } else if (router_scores && !o.score_all) {
        KNN_TRY(idx_router_first(router_scores, Q, nb, ix->block_count.as<int>(), w.lb.as<float>(), s),
                "router first probe");
    }
    if (timing) KNN_TRY(hipEventRecord(evr[1], s), "event");
    KNN_TRY(idx_pick_blocks(score, Q, nb, P, w.lb.as<float>(), ix->block_count.as<int>(),
                            w.chosen.as<int>(), w.lb_rest.as<float>(), s),
            "pick blocks");
    // every scan threshold starts at +inf, or (one-pass ball probes with
    // key-form scans) is seeded by the ball selection, which writes them all
    if (!one_pass || nch_res) KNN_TRY(hipMemsetD32Async((hipDeviceptr_t)w.theta.p, 0xFF800000u, Q, s), "memset");
    if (one_pass) {
        // keys ||r||^2 - 2 r.q (not residual keys): thresholds seeded from each query's reach
        double eq = 0.0, er = 0.0;
        if (nch_bf) {
            bf3_error_bound(dim, nch_bf, &eq, &er);
        } else {
            eq = 2.0 * (dim + 4) * 0x1p-24;
            er = 3.1 * eq;
        }
        KNN_TRY(idx_ball_select(w.qpm.as<float>(), pm_ld, dim, Q, kp, w.chosen.as<int>(), P, ix->fcent.as<float>(),
                                ix->frad.as<double>(), ix->ball_cnt.as<int>(), ix->fb_start.as<int>(),
                                w.bchosen.as<int>(), Pb, P * ix->max_bpb, w.lb_rest.as<float>(), s,
                                w.qnorm.as<float>(), ix->rmax.as<unsigned>(), eq, er,
                                nch_res ? nullptr : w.theta.as<unsigned>()),
                "balls");
    }
    else if (balls)
        KNN_TRY(idx_ball_first(w.qpm.as<float>(), pm_ld, dim, Q, w.chosen.as<int>(), P, ix->fcent.as<float>(),
                               ix->frad.as<double>(), ix->fb_start.as<int>(), w.bchosen.as<int>(), Pb, s),
                "first balls");
    if (timing) KNN_TRY(hipEventRecord(ev[1], s), "event");
    if (timing) KNN_TRY(hipEventRecord(ev[2], s), "event");
    // the scan units each query visits: Pb per query
    int *units = balls ? w.bchosen.as<int>() : w.chosen.as<int>();
    int tbase = 0;
    for (int ph = 0; ph < nphase; ++ph) {
        if (balls && ph == 1) {
            // the other balls, now that phase 0 has set the thresholds
            double eq = 0.0, er = 0.0;
            if (nch_bf) {
                bf3_error_bound(dim, nch_bf, &eq, &er);
            } else {
                eq = 2.0 * (dim + 4) * 0x1p-24;
                er = 3.1 * eq;
            }
            KNN_TRY(idx_ball_rest(w.qpm.as<float>(), pm_ld, dim, Q, w.qnorm.as<float>(), w.theta.as<unsigned>(),
                                  ix->rmax.as<unsigned>(), eq, er, nch_res ? 1 : 0, w.chosen.as<int>(), P,
                                  ix->fcent.as<float>(), ix->frad.as<double>(), ix->fb_start.as<int>(), units, Pb,
                                  cap, w.lb_rest.as<float>(), s),
                    "other balls");
        }
        // (query, unit) pairs of this phase -> per-unit query lists -> tasks
        const int base = tbase * qpt;
        KNN_TRY(hipMemsetAsync(w.pcount.p, 0, (size_t)nu * 4, s), "memset");
        KNN_TRY(idx_pair_count(units, Q, Pb, pr[ph], pr[ph + 1], nu, w.pcount.as<int>(), s), "pairs");
        KNN_TRY(idx_unit_tasks(w.pcount.as<int>(), nu, qpt, unit_start, unit_count, w.qstart.as<int>(),
                               w.cursor.as<int>(), w.tasks.as<ScanTask>() + tbase, max_t[ph], w.fb_count.as<int>(), s),
                "tasks");
        KNN_TRY(idx_pair_scatter(units, Q, Pb, pr[ph], pr[ph + 1], nu, w.cursor.as<int>(), w.qlist.as<int>() + base,
                                 w.rows.as<int>(), base, s),
                "pairs");
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
        const int waves = (double)Q * (pr[ph + 1] - pr[ph]) / nu >= kOneWaveQueries || balls ? 1 : 4;
        if (scan_timing) KNN_TRY(hipEventRecord(evk[ph][0], s), "event");
        if (nch_res) {
            sa.packed_res = ix->res_tiles.as<unsigned short>();
            sa.res_norm = ix->res_norm.as<float>();
            sa.res_cent = ix->res_cent.as<float>();
            sa.res_rho = ix->res_rho.as<float>();
            sa.res_row = w.resrow.as<float>() + (size_t)base * 4;
            sa.res_eta_coef = res_eta_coeff(dim, nch_res);
            sa.res_norm_up = res_norm_up(dim);
            sa.res_norm_dn = res_norm_dn(dim);
            KNN_TRY(launch_scan_res_task(sa, nch_res, waves, scan_lds_bytes(kp, waves), s,
                                         waves == 1 ? kTasksPerGroup : 1, merge),
                    "scan kernel launch");
        } else if (nch_bf) {
            sa.packed_bf = ix->packed_bf.as<unsigned short>();
            // one-wave tasks go kTasksPerGroup to a workgroup: a block's
            // consecutive query groups then read its tiles through one CU
            KNN_TRY(launch_scan_bf3_task(sa, nch_bf, waves, scan_lds_bytes(kp, waves), s,
                                         waves == 1 ? kTasksPerGroup : 1, merge),
                    "scan kernel launch");
        } else {
            KNN_TRY(launch_scan(sa, ks, bmode, waves, s), "scan kernel launch");
        }
        if (scan_timing) KNN_TRY(hipEventRecord(evk[ph][1], s), "event");
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
    fa.nrows = Pb;
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
    if (nch_bf) bf3_error_bound(dim, nch_bf, &fa.err_q, &fa.err_r);
    if (nch_res) fa.res_row = w.resrow.as<float>();
    fa.force_fallback = o.force_fallback;
    fa.out_dist = out_dist;
    fa.out_idx = out_idx;
    fa.fb_count = w.fb_count.as<int>();
    fa.fb_list = w.fb_list.as<int>();
    KNN_TRY(launch_select(fa, s), "select kernel launch");
    if (timing) KNN_TRY(hipEventRecord(ev[4], s), "event");
    // queries whose certificate failed: once more through the index with
    // more probes, then (or at once, on a retry) the plain full scan
    // (the count lands in pinned host memory: one DMA, no staging copy)
    static int *nfail_host = nullptr;
    if (!nfail_host) KNN_TRY(hipHostMalloc((void **)&nfail_host, sizeof(int), hipHostMallocPortable), "host alloc");
    KNN_TRY(hipMemcpyAsync(nfail_host, w.fb_count.p, 4, hipMemcpyDeviceToHost, s), "copy");
    KNN_TRY(hipStreamSynchronize(s), "sync");
    const int nfail = *nfail_host;
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
        // the retry widens both the probe count and the candidate margin, and
        // takes FP32 products (a bfloat16 failure is mostly its wider bound)
        po.margin = std::max(2 * margin, kDefaultMargin);
        po.precision = 0;
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
        // the scan stage is its kernel launches alone; the task lists, threshold set-up and (two
        // passes) the second ball choice between them count as the tasks stage
        float scan_k = 0.f;
        for (int ph = 0; ph < nphase; ++ph) {
            float t = 0.f;
            KNN_TRY(hipEventElapsedTime(&t, evk[ph][0], evk[ph][1]), "event");
            scan_k += t;
        }
        ix->last_ms[1] += ix->last_ms[2] - scan_k;
        ix->last_ms[2] = scan_k;
        for (auto &p : evk)
            for (auto &e : p) (void)hipEventDestroy(e);
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
    } else if (scan_timing) {
        // (the stream was synchronised above, after the select)
        float scan_k = 0.f;
        for (int ph = 0; ph < nphase; ++ph) {
            float t = 0.f;
            KNN_TRY(hipEventElapsedTime(&t, evk[ph][0], evk[ph][1]), "event");
            scan_k += t;
        }
        for (auto &p : evk)
            for (auto &e : p) (void)hipEventDestroy(e);
        ix->last_fail = nfail;
        for (float &m : ix->last_ms) m = -1.f;
        for (float &m : ix->rank_ms) m = -1.f;
        ix->last_ms[2] = scan_k;
        g_stage_ms[1] = scan_k;
    }
    // what knnhip_index_last_work reads back from this search's task list
    if (depth == 0) {
        ix->last_ref_bytes = nch_res ? 32ll * nch_res : nch_bf ? 64ll * nch_bf : 4ll * ix->dpad;
        ix->last_prec = nch_res ? 2 : nch_bf ? 1 : 0;
        ix->last_balls = balls;
        ix->last_rows = Pb;
    }
    g_last_fb = w.fb_count.as<int>();
    g_last_nonfinite = nullptr;
    return true;
}

}  // namespace
}  // namespace knnhip

using namespace knnhip;

extern "C" {

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
    CallOrder order((hipStream_t)stream);
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
    CallOrder order((hipStream_t)stream);
    knnhip_options o = opt ? *opt : default_options(0);
    char buf[256];
    const char *prec = o.precision == 2 && scan_res_nch(ix->dim)           ? "bf16-residual"
                       : o.precision >= 1 && scan_bf3_task_nch(ix->dim) ? "bf16x3"
                                                                        : "fp32";
    snprintf(buf, sizeof(buf),
             "{\"index_blocks\": %d, \"probes\": %d, \"router\": %s, \"scores_rank\": \"%s\", "
             "\"query_nb\": %d, \"k\": %d, \"precision\": \"%s\"}",
             ix->nb, std::min(probes, ix->nb), scores ? "true" : "false",
             !scores ? "none" : o.score_all ? "all" : "first", query_nb, k, prec);
    g_last_plan = buf;
    return index_search(ix, g_iw, 0, query, query_nb, query_ld, k, probes, scores, knn_dist,
                        knn_index, o, (hipStream_t)stream)
               ? 1
               : 0;
}

int knnhip_index_search_routed(knnhip_index *ix, knnhip_router *router, const float *query,
                               int query_nb, long long query_ld, int k, int probes, float *knn_dist,
                               int *knn_index, const knnhip_options *opt, void *stream) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!ix || !router) {
        set_error("null index or router");
        return 0;
    }
    if (!check_args(ix->ref, ix->R, query, query_nb, ix->dim, k, knn_dist, knn_index)) return 0;
    if (query_ld < query_nb || probes < 1) {
        set_error("bad search arguments");
        return 0;
    }
    if (router_dim(router) != ix->dim || router_nblocks(router) != ix->nb) {
        set_error("router (dim %d, %d blocks) does not match the index (dim %d, %d blocks)",
                  router_dim(router), router_nblocks(router), ix->dim, ix->nb);
        return 0;
    }
    if (!ensure_device()) return 0;
    CallOrder order((hipStream_t)stream);
    knnhip_options o = opt ? *opt : default_options(0);
    char buf[256];
    const char *prec = o.precision == 2 && scan_res_nch(ix->dim)           ? "bf16-residual"
                       : o.precision >= 1 && scan_bf3_task_nch(ix->dim) ? "bf16x3"
                                                                        : "fp32";
    snprintf(buf, sizeof(buf),
             "{\"index_blocks\": %d, \"probes\": %d, \"router\": true, \"router_kernel\": \"%s\", "
             "\"scores_rank\": \"%s\", \"query_nb\": %d, \"k\": %d, \"precision\": \"%s\"}",
             ix->nb, std::min(probes, ix->nb), o.score_all ? "scores" : "fused first probe",
             o.score_all ? "all" : "first", query_nb, k, prec);
    g_last_plan = buf;
    return index_search(ix, g_iw, 0, query, query_nb, query_ld, k, probes, nullptr, knn_dist, knn_index,
                        o, (hipStream_t)stream, router)
               ? 1
               : 0;
}

int knnhip_index_bounds(knnhip_index *ix, const float *query, int query_nb, long long query_ld,
                        int precision, float *lb, void *stream) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!ix || !query || !lb || query_nb < 1 || query_ld < query_nb || precision < 0 || precision > 3) {
        set_error("bad bound arguments");
        return 0;
    }
    if (!ensure_device()) return 0;
    CallOrder order((hipStream_t)stream);
    IndexWork &w = g_iw;
    if (!w.need(w.qnorm, (size_t)query_nb * 4) || !w.need(w.qpm, (size_t)query_nb * ix->pm_ld * 4) ||
        !w.need(w.lb, (size_t)lb_elems(query_nb, ix->nb) * 4)) {
        set_error("Memory allocation error");
        return 0;
    }
    const hipStream_t s = (hipStream_t)stream;
    NormArgs nq{};
    nq.p = query;
    nq.ld = query_ld;
    nq.n = query_nb;
    nq.n_pad = query_nb;
    nq.dim = ix->dim;
    nq.oct = (ix->dim + 7) / 8;
    nq.out = w.qnorm.as<float>();
    nq.pm = w.qpm.as<float>();
    nq.pm_ld = ix->pm_ld;
    if (launch_norms(nq, s) != hipSuccess) {
        set_error("norm kernel launch");
        return 0;
    }
    // (the searches' tiled layout, then rows of query_nb)
    if (!index_bounds(ix, w.qpm.as<float>(), w.qnorm.as<float>(), query_nb, (precision & 1) != 0,
                      w.lb.as<float>(), s, (precision & 2) != 0))
        return 0;
    if (idx_lb_untile(w.lb.as<float>(), query_nb, ix->nb, lb, s) != hipSuccess) {
        set_error("bound layout kernel launch");
        return 0;
    }
    return 1;
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
    CallOrder order((hipStream_t)stream);
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

int knnhip_index_last_work(const knnhip_index *ix, long long *exec_pairs, long long *pairs,
                           long long *tile_bytes, long long *unit_bytes, int *precision) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!ix) {
        set_error("null index");
        return 0;
    }
    // the task list of the last search is still in the search workspace
    long long ex = 0, pr = 0, tb = 0, ub = 0;
    const long long nt = std::min<long long>(ix->last_tasks, (long long)(g_iw.tasks.n / sizeof(ScanTask)));
    if (nt > 0) {
        std::vector<ScanTask> tl(nt);
        KNN_TRY(hipMemcpy(tl.data(), g_iw.tasks.p, (size_t)nt * sizeof(ScanTask), hipMemcpyDeviceToHost), "copy");
        // distinct tiles: every 32-reference tile of a unit the search scanned, once
        std::vector<std::pair<int, int>> tiles;
        for (const ScanTask &t : tl) {
            if (t.qcount <= 0) continue;
            const long long steps = (t.ref_end - t.ref_begin + kRefsPerStep - 1) / kRefsPerStep;
            ex += steps * kRefsPerStep * kQueriesPerWave;
            pr += (long long)(t.ref_end - t.ref_lo) * t.qcount;
            tb += steps * kRefsPerStep * ix->last_ref_bytes;
            tiles.emplace_back(t.ref_lo / kRefsPerTile, (t.ref_end + kRefsPerTile - 1) / kRefsPerTile);
        }
        std::sort(tiles.begin(), tiles.end());
        long long covered = 0;
        int end = -1;
        for (const auto &r : tiles) {
            const int lo = std::max(r.first, end);
            if (r.second > lo) covered += r.second - lo;
            end = std::max(end, r.second);
        }
        ub = covered * kRefsPerTile * ix->last_ref_bytes;
    }
    if (exec_pairs) *exec_pairs = ex;
    if (pairs) *pairs = pr;
    if (tile_bytes) *tile_bytes = tb;
    if (unit_bytes) *unit_bytes = ub;
    if (precision) *precision = ix->last_prec;
    return 1;
}

int knnhip_index_last_units(const knnhip_index *ix, int *balls, int *rows_per_query, int *nballs) {
    if (!ix) {
        set_error("null index");
        return 0;
    }
    if (balls) *balls = ix->last_balls ? 1 : 0;
    if (rows_per_query) *rows_per_query = ix->last_rows;
    if (nballs) *nballs = ix->ball_ranges ? ix->nf : 0;
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
    CallOrder order((hipStream_t)stream);
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
}  // extern "C"
