/*
 * knncuda.h — drop-in C/C++ interface of the MI355X exact k-NN library
 * (libknnhip.so, built from knn-cuda_amd/csrc).
 *
 * Point sets are dim x n matrices stored row-major: coordinate d of point i is
 * at ptr[d*n + i].  Results are k x query_nb row-major: neighbour rank j of
 * query i is at knn_dist[j*query_nb + i] / knn_index[j*query_nb + i], ranks in
 * ascending distance, indexes 0-based, equal distances ordered by the lower
 * reference index.  Every entry point returns the ground-truth result of the
 * reference (code/test.cpp:129-175): distances are sqrtf of the FP32 sum of
 * squared differences accumulated in ascending d, bit for bit.
 */
#ifndef KNNHIP_KNNCUDA_H
#define KNNHIP_KNNCUDA_H

#include <stddef.h>

/* The library is built with hidden visibility: what this header declares is
 * all it exports. */
#pragma GCC visibility push(default)

#ifdef __cplusplus
/*
 * The three reference entry points, with C++ linkage so that code compiled
 * against the reference header links against this library unchanged.
 * Host pointers in and out; the call returns once the results are in host
 * memory.  On failure: a line starting "ERROR:" on stdout and false.
 *
 * Replaces code/knncuda.h:14-21 (knn_cuda_global).
 */
bool knn_cuda_global(const float *ref, int ref_nb, const float *query, int query_nb, int dim,
                     int k, float *knn_dist, int *knn_index);
/* Replaces code/knncuda.h:37-44 (knn_cuda_texture).  No texture objects and no
 * width limit: reference tiles are staged through LDS. */
bool knn_cuda_texture(const float *ref, int ref_nb, const float *query, int query_nb, int dim,
                      int k, float *knn_dist, int *knn_index);
/* Replaces code/knncuda.h:61-68 (knn_cublas).  The ||r||^2 - 2 q.r contraction
 * runs on the library's own MFMA kernel (no vendor BLAS); distances are never
 * NaN. */
bool knn_cublas(const float *ref, int ref_nb, const float *query, int query_nb, int dim, int k,
                float *knn_dist, int *knn_index);
extern "C" {
#endif

/* C-linkage aliases of the three entry points above, for FFI callers
 * (ctypes, cgo, JNI).  Return 1 on success, 0 on failure. */
int knnhip_global(const float *ref, int ref_nb, const float *query, int query_nb, int dim, int k,
                  float *knn_dist, int *knn_index);
int knnhip_texture(const float *ref, int ref_nb, const float *query, int query_nb, int dim, int k,
                   float *knn_dist, int *knn_index);
int knnhip_cublas(const float *ref, int ref_nb, const float *query, int query_nb, int dim, int k,
                  float *knn_dist, int *knn_index);
/* Distance to the k-th nearest neighbour of every query (query_nb floats,
 * host pointers), as the reference's legacy programs without indexes return
 * it (legacy/KNN.CUDA.without.indexes/knn_cuda_without_indexes.cu:278,
 * knn(ref, ref_width, query, query_width, height, k, dist)): equal to row k-1
 * of knn_cuda_global's distances. */
int knnhip_kth_distance(const float *ref, int ref_nb, const float *query, int query_nb, int dim,
                        int k, float *kth_dist);

/* Options of the extended entry point.  Zero-initialise, then set fields;
 * zero means "library default" for every field. */
typedef struct knnhip_options {
    int variant;        /* a label only: 0 global (default), 1 texture, 2 cublas.  The three
                           reference variants compute one result, so all three run the same
                           kernels (recorded as "variant" in knnhip_last_plan()) */
    int margin;         /* extra candidates kept per query beyond k (default 16) */
    int slices;         /* reference slices scanned in parallel (default: chosen) */
    int force_fallback; /* 1: every query takes the exact full-scan path (testing) */
    int timing;         /* 1: record per-stage device times (knnhip_last_stage_ms); 2 (index
                           searches): the scan stage's alone, the others read -1 */
    int precision;      /* 0: FP32 products; 1: bfloat16 three-term products (same exact
                           result; dim <= 192 and k + margin <= 107, else FP32); 2: index
                           searches take single-pass bfloat16 products on residuals against
                           the centre of each scanned block or ball (dim <= 256, else as 1),
                           plain searches as 1 */
    int block_queries;  /* FP32 scan shape: 0 chosen, 32 (reference tiles per wave), 128
                           (reference tiles shared through LDS; dim <= 192, k + margin <= 107) */
    int score_all;      /* index searches given scores: 0 the scores choose each query's
                           first probe (scanned first), the lower bounds rank the others;
                           1 the scores rank every probe */
    int probe_units;    /* index searches, what the scan visits in the probed blocks (indexes
                           that know their balls' position ranges: built ones and version-2
                           files; others always take 1): 0 the balls that may hold a neighbour,
                           chosen in one pass from the balls' reach; 1 whole blocks; 2 balls in
                           two passes (the first block's nearest ball, then the balls below the
                           threshold it left) */
    int reserved[4];
} knnhip_options;

/*
 * Extended entry point: same problem as knn_cuda_global, with every pointer a
 * device pointer (ref is dim rows of ref_ld floats, query dim rows of query_ld
 * floats; outputs k x query_nb, leading dimension query_nb).  Work is queued
 * on `stream` (a hipStream_t, NULL for the null stream).  While a wider
 * candidate list is left to retry with (margin below 256, the usual case)
 * the call synchronises the stream once, after the re-rank, to read how many
 * queries failed their certificate; the work queued after that (the retry's
 * scatter, the exact per-query scan) may still be running when it returns.
 * With margin 256 or more nothing is read back and the call returns without
 * synchronising.  Calls are ordered on the device: a call on another stream
 * than the previous library call's waits for that call's work first (the
 * library's device buffers are shared).  Inputs must be finite: a point whose
 * squared norm is NaN or infinite makes every output index -1 and distance NaN,
 * and knnhip_last_fallback_count() then returns -1.  Returns 1 on success, 0
 * on failure (knnhip_last_error).  No counterpart in the reference: the
 * reference only takes host pointers (code/knncuda.h:14-21).
 */
int knnhip_search_device(const float *ref, int ref_nb, long long ref_ld, const float *query,
                         int query_nb, long long query_ld, int dim, int k, float *knn_dist,
                         int *knn_index, const knnhip_options *opt, void *stream);

/*
 * knnhip_search_device for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments, same exact
 * result; the call makes no hipMalloc / hipFree and calls no allocating
 * routine: every buffer besides the caller's inputs and outputs lives in one
 * static scratch array of 64 MiB per device (held from the library's first
 * use on that device until the process ends).  A search whose buffers do not
 * fit (about 2 * 4 * ref_nb * dim bytes of reference layouts plus heaps and
 * query rows) returns 0 with an error naming the size.  Queries that fail the
 * candidate certificate go straight to the exact per-query scan (no
 * wider-margin rescan, whose buffers would be allocated); options.timing is
 * ignored.  Every call on a device shares that device's scratch: calls that
 * overlap on one device race, so callers order them (the library orders its
 * own calls on the device, as for knnhip_search_device).
 */
int knnhip_search_device_noalloc(const float *ref, int ref_nb, long long ref_ld, const float *query,
                                 int query_nb, long long query_ld, int dim, int k, float *knn_dist,
                                 int *knn_index, const knnhip_options *opt, void *stream);

/* Number of queries of the last call that failed the candidate certificate
 * and were answered by the exact full scan.  Synchronises the last stream. */
int knnhip_last_fallback_count(void);

/* Device time in milliseconds of stage `stage` in the last call made with
 * options.timing = 1 (0 norms, 1 scan, 2 select/re-rank, 3 fallback); -1 if
 * not recorded. */
float knnhip_last_stage_ms(int stage);

/* Text of the last error reported by this library (empty if none). */
const char *knnhip_last_error(void);

/* Launch plan of the last search as a JSON object (slices, heap size, blocks
 * per CU, ...), for benchmarks and diagnostics. */
const char *knnhip_last_plan(void);

/* Seeded synthetic point sets written directly to device memory, dim x n
 * row-major.  Uniform: scale * U[0,1) from a counter-based hash (bit-identical
 * to the CPU generator in the oracle).  Mixture: ncomp isotropic Gaussians of
 * standard deviation sigma, centres uniform in [-1,1]^dim drawn from
 * centre_seed. */
int knnhip_generate_uniform(float *out, int n, int dim, unsigned long long seed, float scale,
                            void *stream);
int knnhip_generate_mixture(float *out, int n, int dim, int ncomp, float sigma,
                            unsigned long long centre_seed, unsigned long long point_seed,
                            void *stream);

/*
 * Block index (router form).  No counterpart in the reference, which only
 * has the brute-force scan (code/knncuda.h).  The index covers the reference
 * set with 4 * nblocks k-means balls (`iters` Lloyd iterations) grouped into
 * `nblocks` blocks of balanced size, and keeps a block-contiguous copy; a
 * search scans each query's `probes` highest-ranked blocks and skips the
 * rest only where the triangle inequality (over the blocks' balls) proves
 * they hold no neighbour, so results equal knn_cuda_global's exactly.
 * `scores` (nblocks x query_nb, row-major, device: one column per query, like
 * the point layout), e.g. router outputs, picks each query's first probe, the
 * block scanned first (all probes with options.score_all); the distance lower
 * bounds rank the rest, and all of them when scores is NULL.  The index keeps a pointer to `ref`
 * (the full-scan fallback reads it): the caller keeps that buffer alive and
 * unchanged.  nblocks <= min(ref_nb, 32768).
 */
typedef struct knnhip_index knnhip_index;
int knnhip_index_build(const float *ref, int ref_nb, long long ref_ld, int dim, int nblocks,
                       int iters, knnhip_index **out, void *stream);
int knnhip_index_search(knnhip_index *index, const float *query, int query_nb, long long query_ld,
                        int k, int probes, const float *scores, float *knn_dist, int *knn_index,
                        const knnhip_options *opt, void *stream);
void knnhip_index_free(knnhip_index *index);
int knnhip_index_info(const knnhip_index *index, int *nblocks, int *n_pos, float *build_ms);
/* The lower bounds a search ranks and prunes by: lb[b * query_nb + q] (device,
 * nblocks x query_nb) <= the distance from query q to every point of block b
 * (+inf for an empty block).  precision bit 0 (1): bfloat16 three-term
 * products (dim <= 128), as searches with options.precision = 1 use; bit 1
 * (2): the per-ball fold form of the same bounds instead of the slot form
 * that searches use (they agree bit for bit).  For tests and diagnostics. */
int knnhip_index_bounds(knnhip_index *index, const float *query, int query_nb, long long query_ld,
                        int precision, float *lb, void *stream);
/* Versioned file of an index's build (block layout, permutation, balls; see
 * knn_engine.hip for the byte layout).  Loading rebuilds the tiles from the
 * same reference set (ref_nb and dim must match the file). */
int knnhip_index_save(const knnhip_index *index, const char *path);
int knnhip_index_load(const char *path, const float *ref, int ref_nb, long long ref_ld, int dim,
                      knnhip_index **out, void *stream);
/* Device times (ms) of the last search made with options.timing = 1: query
 * norms, block bounds, block choice, task build, scan, select/re-rank,
 * full-scan fallback (n <= 7 values written). */
int knnhip_index_stage_ms(const knnhip_index *index, float *ms, int n);
/* Scan work of the last index search (read back from its task list, so call
 * it before the next index search): the (reference, query) pairs its MFMA
 * tiles executed (whole 128-reference steps times 32 query columns per task),
 * the pairs its tasks held, the bytes of reference tiles its tasks read, the
 * bytes of the distinct tiles they cover (each scan unit's 32-reference tiles
 * once: what a scan that read every unit once would read), and the products
 * it used (0 FP32, 1 bfloat16 three-term, 2 bfloat16 on centre residuals).
 * Any pointer may be NULL.  For benchmarks (roofline figures). */
int knnhip_index_last_work(const knnhip_index *index, long long *exec_pairs, long long *pairs,
                           long long *tile_bytes, long long *unit_bytes, int *precision);
/* Scan units of the last index search: balls (1) or whole blocks (0), its
 * heap rows per query (units visited at most), and the index's ball count
 * when it knows the balls' position ranges (else 0).  Any pointer may be
 * NULL.  No counterpart in the reference (it has no index). */
int knnhip_index_last_units(const knnhip_index *index, int *balls, int *rows_per_query, int *nballs);
/* Block layout of the index: block_start / block_count (host, nblocks ints
 * each: first position and size of each block in the block-contiguous copy)
 * and ref_block (device, ref_nb ints: the block holding each reference point,
 * written on `stream`).  Any of the three may be NULL. */
int knnhip_index_layout(const knnhip_index *index, int *block_start, int *block_count,
                        int *ref_block, void *stream);

/*
 * Learned router for the block index.  No counterpart in the reference.  An
 * MLP dim -> hidden -> hidden -> nblocks (ReLU after both hidden layers,
 * bfloat16 products with FP32 accumulation) scores the blocks of each query;
 * its scores go to knnhip_index_search as `scores`.  Training runs on the
 * GPU: exact neighbours of the training queries (dim-major, like the point
 * sets) come from the plain search, the target of a query is the share of
 * its k neighbours in each block, the loss is softmax cross-entropy and the
 * optimiser Adam (beta1 0.9, beta2 0.999, eps 1e-8).  Parameters are one FP32
 * vector: W1 [hidden][dim], b1, W2 [hidden][hidden], b2, W3 [nblocks][hidden],
 * b3 (weights row-major, out x in).  Device pointers unless said otherwise.
 */
typedef struct knnhip_router knnhip_router;
int knnhip_router_create(int dim, int hidden, int nblocks, unsigned long long seed,
                         knnhip_router **out);
void knnhip_router_free(knnhip_router *router);
long long knnhip_router_num_params(const knnhip_router *router);
/* Copy the parameter vector to (set = 0) or from (set = 1) host memory; a set
 * also resets the optimiser state. */
int knnhip_router_params(knnhip_router *router, float *host_params, long long n, int set);
/* One forward + backward pass over n samples x (dim x n, leading dimension
 * ld) with targets tgt (k x n, leading dimension ldt: the block of each
 * neighbour): gradient of the mean loss into grad (num_params floats, may be
 * NULL) and the mean loss into *loss (host, may be NULL). */
int knnhip_router_loss_grad(knnhip_router *router, const float *x, int n, long long ld,
                            const int *tgt, long long ldt, int k, float *grad, float *loss,
                            void *stream);
/* The same pass followed by one Adam update with learning rate lr. */
int knnhip_router_step(knnhip_router *router, const float *x, int n, long long ld, const int *tgt,
                       long long ldt, int k, float lr, void *stream);
/* Train on n queries (dim x n, ld) against the index's blocks: k exact
 * neighbours per query, `epochs` shuffled passes in batches of `batch`.
 * *final_loss (host, may be NULL) gets the mean loss of the last epoch. */
int knnhip_router_train(knnhip_router *router, const knnhip_index *index, const float *train_query,
                        int n, long long ld, int k, int epochs, int batch, float lr,
                        unsigned long long seed, float *final_loss, void *stream);
/* Versioned file of a router's parameters (dimensions + FP32 vector). */
int knnhip_router_save(const knnhip_router *router, const char *path);
int knnhip_router_load(const char *path, knnhip_router **out);
/* Block scores of query_nb queries into scores (nblocks x query_nb). */
int knnhip_router_scores(knnhip_router *router, const float *query, int query_nb,
                         long long query_ld, float *scores, void *stream);
/* First probe of each query: the block with the highest score, from one fused
 * kernel (the three layers run back to back per 64 queries, bfloat16
 * products, and only the argmax leaves the kernel; ties keep the lower
 * block).  top1: query_nb ints on the device.  Needs hidden 64, 128 or 256
 * and dim <= 512.  No counterpart in the reference. */
int knnhip_router_top1(knnhip_router *router, const float *query, int query_nb, long long query_ld,
                       int *top1, void *stream);
/* knnhip_index_search with the router's first probe for each query taken
 * from the fused kernel above (options.score_all = 1: the router's scores
 * rank every probe).  The router must match the index (dim, nblocks). */
int knnhip_index_search_routed(knnhip_index *index, knnhip_router *router, const float *query,
                               int query_nb, long long query_ld, int k, int probes, float *knn_dist,
                               int *knn_index, const knnhip_options *opt, void *stream);

/*
 * Batched form.  No counterpart in the reference.  Runs many independent
 * searches, each exactly as knn_cuda_global would (device pointers: dim-major
 * point sets with leading dimensions, k x query_nb outputs).  Problems
 * sharing (dim, k) form one class and run as one scan launch; the references
 * per scan task (the class's tile shape) come from the tuned table, or from
 * a cost model when the class is untuned.  knnhip_batched_autotune times the
 * candidate shapes on the given problems and records the fastest per class;
 * knnhip_batched_table writes the table as JSON (returns the length needed,
 * terminating zero included); knnhip_batched_set_slice sets or (0) clears
 * an entry.  The search synchronises the stream before it returns.
 */
typedef struct knnhip_problem {
    const float *ref;
    int ref_nb;
    long long ref_ld;
    const float *query;
    int query_nb;
    long long query_ld;
    int dim;
    int k;
    float *knn_dist;
    int *knn_index;
} knnhip_problem;
int knnhip_search_batched(const knnhip_problem *problems, int count, const knnhip_options *opt,
                          void *stream);
/* Certificate failures of the last batched search (answered by the plain
 * search of their problem, so still exact). */
long long knnhip_batched_last_failures(void);
int knnhip_batched_autotune(const knnhip_problem *problems, int count, int reps, void *stream);
int knnhip_batched_table(char *buf, int len);
int knnhip_batched_set_slice(int dim, int k, int slice_len);

/*
 * Device memory budget of the host-pointer entry points (knn_cuda_global,
 * knn_cuda_texture, knn_cublas, their knnhip_ aliases, knnhip_kth_distance).
 * A search whose buffers would exceed it streams: the reference set goes
 * through two chunk buffers (the next chunk's copy overlapping the current
 * chunk's search), the queries go in groups, and the per-chunk results are
 * merged on the device, with the same exact result.  0 (the default) means
 * 90 % of the free device memory plus this library's caches, the share the
 * reference's legacy memory-adaptive split uses
 * (legacy/KNN.CUDA.with.indexes/knn_cuda_with_indexes.cu:49,340-349, which
 * splits only the queries).  A streamed call frees the library's cached
 * buffers first.  Returns 1, or 0 for a negative value.
 */
int knnhip_set_device_budget(long long bytes);
long long knnhip_device_budget(void);
/* Reference chunk and query group sizes a streamed search of these sizes
 * would use within device_bytes (ref_chunk == ref_nb: references resident).
 * Returns 0 when nothing fits (knnhip_last_error says why). */
int knnhip_stream_chunks(int ref_nb, int query_nb, int dim, int k, long long device_bytes,
                         int *ref_chunk, int *query_chunk);

/* Library version string. */
const char *knnhip_version(void);

#ifdef __cplusplus
}
#endif

#pragma GCC visibility pop

#endif /* KNNHIP_KNNCUDA_H */
