// Host side of the learned router: parameters, forward / backward / Adam
// steps on the GPU, training on exact neighbours' blocks, scoring, and the
// router file format.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "knn_engine.h"
#include "knn_index.h"
#include "knn_router.h"

// The opaque router handle of the C API: MLP dim -> H -> H -> nb.
struct knnhip_router {
    int dim = 0, H = 0, nb = 0;
    knnhip::RouterLayout lay{};
    knnhip::DevBuf params, grad, m, v;
    int step = 0;
    // batch workspace (grown on demand)
    knnhip::DevBuf h1, h2, z3, dz3, d2, d1, loss, lsum;
    // bfloat16 weights of the fused inference kernel, converted again
    // whenever the parameters have changed (gen counts the changes)
    knnhip::DevBuf wbf;
    int gen = 0, wbf_gen = -1;
    knnhip::DevBuf scores;  // score matrix where the fused kernel does not apply
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
    r->gen += 1;
    KNN_TRY(launch_adam(r->params.as<float>(), r->grad.as<float>(), r->m.as<float>(),
                        r->v.as<float>(), r->lay.total, lr, kAdamB1, kAdamB2, kAdamEps, r->step, s),
            "adam");
    return true;
}

}  // namespace

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

int router_dim(const knnhip_router *r) { return r->dim; }
int router_nblocks(const knnhip_router *r) { return r->nb; }

// First probe of each of Q queries (dim-major, ld): the non-empty block
// (counts[b] != 0; counts may be null) with the highest score, into top1
// and/or as bound (top1, q) = -inf of the tiled bound matrix lb (lb_at).
// One fused kernel where the router's shape allows it (hidden 64, 128 or
// 256; dim up to 512), else the score matrix and the index's first-probe
// kernel (lb only).
bool router_pick_first(knnhip_router *r, const float *q, int Q, long long ld, const int *counts,
                       int *top1, float *lb, hipStream_t s) {
    const int dpad = (r->dim + 15) / 16 * 16;
    if (router_first_supported(dpad, r->H)) {
        if (r->wbf_gen != r->gen) {
            const size_t n = (size_t)r->H * dpad + (size_t)r->H * r->H + (size_t)r->nb * r->H;
            if (r->wbf.n < n * 2 && !r->wbf.alloc(n * 2)) {
                set_error("Memory allocation error");
                return false;
            }
            KNN_TRY(launch_router_bf16_weights(r->params.as<float>(), r->lay, r->dim, dpad, r->H, r->nb,
                                               r->wbf.as<unsigned short>(), s),
                    "router weights");
            r->wbf_gen = r->gen;
        }
        const float *w = r->params.as<float>();
        const unsigned short *wb = r->wbf.as<unsigned short>();
        RouterFirstArgs a{};
        a.x = q;
        a.ldx = ld;
        a.Q = Q;
        a.dim = r->dim;
        a.dpad = dpad;
        a.w1 = wb;
        a.w2 = wb + (size_t)r->H * dpad;
        a.w3 = a.w2 + (size_t)r->H * r->H;
        a.b1 = w + r->lay.b_off[0];
        a.b2 = w + r->lay.b_off[1];
        a.b3 = w + r->lay.b_off[2];
        a.nb = r->nb;
        a.counts = counts;
        a.top1 = top1;
        a.lb = lb;
        KNN_TRY(launch_router_first_fused(a, r->H, s), "router first probe");
        return true;
    }
    if (top1) {
        set_error("router shape (dim %d, hidden %d) has no fused first-probe kernel", r->dim, r->H);
        return false;
    }
    if (!counts) {
        set_error("block counts needed");
        return false;
    }
    if (r->scores.n < (size_t)r->nb * Q * 4 && !r->scores.alloc((size_t)r->nb * Q * 4)) {
        set_error("Memory allocation error");
        return false;
    }
    if (!router_scores(r, q, Q, ld, r->scores.as<float>(), s)) return false;
    KNN_TRY(idx_router_first(r->scores.as<float>(), Q, r->nb, counts, lb, s), "router first probe");
    return true;
}

namespace {

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

using namespace knnhip;

extern "C" {

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
        r->gen += 1;
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
    CallOrder order(s);
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
    CallOrder order((hipStream_t)stream);
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
    CallOrder order((hipStream_t)stream);
    return router_train(r, ix, train_query, n, ld, k, epochs, std::min(batch, n), lr, seed,
                        final_loss, (hipStream_t)stream)
               ? 1
               : 0;
}

int knnhip_router_top1(knnhip_router *r, const float *query, int query_nb, long long query_ld,
                       int *top1, void *stream) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_last_error.clear();
    if (!r || !query || !top1 || query_nb < 1 || query_ld < query_nb) {
        set_error("bad router top1 arguments");
        return 0;
    }
    if (!ensure_device()) return 0;
    CallOrder order((hipStream_t)stream);
    return router_pick_first(r, query, query_nb, query_ld, nullptr, top1, nullptr, (hipStream_t)stream) ? 1
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
    CallOrder order((hipStream_t)stream);
    return router_scores(r, query, query_nb, query_ld, scores, (hipStream_t)stream) ? 1 : 0;
}
}  // extern "C"
