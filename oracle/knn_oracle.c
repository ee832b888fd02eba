/*
 * CPU oracle for exact brute-force k-NN.  TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker, never the product: only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py load it.  The HIP
 * library in knn-cuda_amd/ never links, loads or calls anything here.
 *
 * Built with -O2 -ffp-contract=off and without -ffast-math (oracle/Makefile):
 * a fused multiply-add in the distance loop would change the bits of the
 * result, and the restatement must reproduce the reference bit for bit.
 *
 * Contents
 *   oracle_knn_c        restatement of the reference ground truth knn_c
 *                       (code/test.cpp:129-175, distance :54-67, top-k :85-114)
 *   oracle_knn_c_omp    the same computation with queries spread over threads
 *                       (each query is computed exactly as in oracle_knn_c, so
 *                       the results are bit-identical)
 *   oracle_knn_f64      FP64 brute force (true neighbours, for error statements)
 *   oracle_canonical_sq the FP32 squared-distance sum of one (ref, query) pair
 *   oracle_check        the reference accuracy check (code/test.cpp:216-272)
 *   oracle_gen_*        seeded counter-based generators shared with the GPU
 *
 * Pinning (see tests/test_oracle.py): against golden vectors written by the
 * reference's own knn_c (tests/golden/), against the compiled reference in
 * oracle/_ref when present, and against closed forms (FP64 brute force on
 * small inputs, points on a line, duplicates, k=1, k=ref_nb).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Squared Euclidean distance exactly as the reference ground truth forms it:
 * d ascending, difference rounded to FP32, square rounded to FP32, running sum
 * rounded to FP32 after every add (code/test.cpp:61-66, before the sqrtf). */
float oracle_canonical_sq(const float *ref, int ref_nb, const float *query, int query_nb,
                          int dim, int ref_index, int query_index) {
    float acc = 0.0f;
    const float *r = ref + ref_index;
    const float *q = query + query_index;
    for (int d = 0; d < dim; ++d) {
        float t = r[(size_t)d * (size_t)ref_nb] - q[(size_t)d * (size_t)query_nb];
        float sq = t * t;
        acc = acc + sq;
    }
    return acc;
}

/* The ground-truth key of the reference is sqrtf of the sum above
 * (code/test.cpp:66). */
static void topk_insert_block(const float *keys, int n, int j0, int k, int *filled_io,
                              float *best_key, int *best_idx);

/*
 * Top-k selection with the semantics of the reference's modified insertion
 * sort (code/test.cpp:85-114): candidates arrive in increasing index order; a
 * candidate is dropped when the list already holds k entries and its key is
 * not strictly below the current k-th key (:98); otherwise it is placed after
 * every entry whose key is <= its own (the strict '>' shift at :104).  The
 * result is therefore the k smallest (key, index) pairs in lexicographic
 * order.  Written as a bounded sorted list rather than in place over the full
 * distance array; the output is the same.
 *
 * The keys are formed kBlock references at a time: for each d, the partial
 * sums of the block's references take their d-th term.  Every reference's sum
 * still adds its terms one by one in ascending d with the same FP32
 * roundings as oracle_canonical_sq (so each key is bit-identical to sqrtf of
 * it), but the dim-major rows are read contiguously instead of one element
 * per row.
 */
enum { kBlock = 256 };
static void topk_one_query(const float *ref, int ref_nb, const float *query, int query_nb,
                           int dim, int k, int qi, float *best_key, int *best_idx) {
    int filled = 0;
    float acc[kBlock];
    for (int j0 = 0; j0 < ref_nb; j0 += kBlock) {
        const int nb = ref_nb - j0 < kBlock ? ref_nb - j0 : kBlock;
        for (int jj = 0; jj < nb; ++jj)
            acc[jj] = 0.0f;
        for (int d = 0; d < dim; ++d) {
            const float qd = query[(size_t)d * (size_t)query_nb + (size_t)qi];
            const float *row = ref + (size_t)d * (size_t)ref_nb + (size_t)j0;
            for (int jj = 0; jj < nb; ++jj) {
                float t = row[jj] - qd;
                float sq = t * t;
                acc[jj] = acc[jj] + sq;
            }
        }
        for (int jj = 0; jj < nb; ++jj)
            acc[jj] = sqrtf(acc[jj]);
        topk_insert_block(acc, nb, j0, k, &filled, best_key, best_idx);
    }
}

/* Insertion of one block of keys (references j0 .. j0 + n - 1, in order). */
static void topk_insert_block(const float *keys, int n, int j0, int k, int *filled_io,
                              float *best_key, int *best_idx) {
    int filled = *filled_io;
    for (int jj = 0; jj < n; ++jj) {
        const int j = j0 + jj;
        float key = keys[jj];
        if (filled == k && !(key < best_key[k - 1]))
            continue;
        int pos = (filled < k) ? filled : k - 1;
        while (pos > 0 && best_key[pos - 1] > key) {
            best_key[pos] = best_key[pos - 1];
            best_idx[pos] = best_idx[pos - 1];
            --pos;
        }
        best_key[pos] = key;
        best_idx[pos] = j;
        if (filled < k)
            ++filled;
    }
    *filled_io = filled;
}

/* Output layout of the reference: rank j of query i at [j*query_nb + i]
 * (code/test.cpp:163-166).  Returns 1 on success, 0 on bad arguments or
 * allocation failure. */
static int knn_c_range(const float *ref, int ref_nb, const float *query, int query_nb,
                       int dim, int k, float *knn_dist, int *knn_index, int q_begin,
                       int q_end) {
    float *key = (float *)malloc((size_t)k * sizeof(float));
    int *idx = (int *)malloc((size_t)k * sizeof(int));
    if (!key || !idx) {
        free(key);
        free(idx);
        return 0;
    }
    for (int i = q_begin; i < q_end; ++i) {
        topk_one_query(ref, ref_nb, query, query_nb, dim, k, i, key, idx);
        for (int j = 0; j < k; ++j) {
            knn_dist[(size_t)j * (size_t)query_nb + (size_t)i] = key[j];
            knn_index[(size_t)j * (size_t)query_nb + (size_t)i] = idx[j];
        }
    }
    free(key);
    free(idx);
    return 1;
}

int oracle_knn_c(const float *ref, int ref_nb, const float *query, int query_nb, int dim,
                 int k, float *knn_dist, int *knn_index) {
    if (ref_nb < 1 || query_nb < 1 || dim < 1 || k < 1 || k > ref_nb)
        return 0;
    return knn_c_range(ref, ref_nb, query, query_nb, dim, k, knn_dist, knn_index, 0, query_nb);
}

/* Subsample form used by the CPU baseline: only queries [q_begin, q_end) are
 * computed; the other output columns are left untouched. */
int oracle_knn_c_queries(const float *ref, int ref_nb, const float *query, int query_nb,
                         int dim, int k, float *knn_dist, int *knn_index, int q_begin,
                         int q_end) {
    if (ref_nb < 1 || query_nb < 1 || dim < 1 || k < 1 || k > ref_nb || q_begin < 0 ||
        q_end > query_nb || q_begin > q_end)
        return 0;
    return knn_c_range(ref, ref_nb, query, query_nb, dim, k, knn_dist, knn_index, q_begin,
                       q_end);
}

int oracle_knn_c_omp(const float *ref, int ref_nb, const float *query, int query_nb,
                     int dim, int k, float *knn_dist, int *knn_index) {
    if (ref_nb < 1 || query_nb < 1 || dim < 1 || k < 1 || k > ref_nb)
        return 0;
    int ok = 1;
#pragma omp parallel reduction(&& : ok)
    {
        float *key = (float *)malloc((size_t)k * sizeof(float));
        int *idx = (int *)malloc((size_t)k * sizeof(int));
        if (!key || !idx) {
            ok = 0;
        } else {
#pragma omp for schedule(dynamic, 4)
            for (int i = 0; i < query_nb; ++i) {
                topk_one_query(ref, ref_nb, query, query_nb, dim, k, i, key, idx);
                for (int j = 0; j < k; ++j) {
                    knn_dist[(size_t)j * (size_t)query_nb + (size_t)i] = key[j];
                    knn_index[(size_t)j * (size_t)query_nb + (size_t)i] = idx[j];
                }
            }
        }
        free(key);
        free(idx);
    }
    return ok;
}

/*
 * FP64 brute force: true Euclidean distance of every pair in double, the k
 * smallest (distance, index) pairs in lexicographic order.  Used to state the
 * error of FP32 results against the real neighbours; its indexes may differ
 * from the FP32 ground truth where FP32 rounding reorders near-ties.
 */
int oracle_knn_f64(const float *ref, int ref_nb, const float *query, int query_nb, int dim,
                   int k, double *knn_dist, int *knn_index) {
    if (ref_nb < 1 || query_nb < 1 || dim < 1 || k < 1 || k > ref_nb)
        return 0;
    int ok = 1;
#pragma omp parallel reduction(&& : ok)
    {
        double *key = (double *)malloc((size_t)k * sizeof(double));
        int *idx = (int *)malloc((size_t)k * sizeof(int));
        if (!key || !idx) {
            ok = 0;
        } else {
#pragma omp for schedule(dynamic, 4)
            for (int i = 0; i < query_nb; ++i) {
                int filled = 0;
                for (int j = 0; j < ref_nb; ++j) {
                    double s = 0.0;
                    for (int d = 0; d < dim; ++d) {
                        double t = (double)ref[(size_t)d * ref_nb + j] -
                                   (double)query[(size_t)d * query_nb + i];
                        s += t * t;
                    }
                    double key_j = sqrt(s);
                    if (filled == k && !(key_j < key[k - 1]))
                        continue;
                    int pos = (filled < k) ? filled : k - 1;
                    while (pos > 0 && key[pos - 1] > key_j) {
                        key[pos] = key[pos - 1];
                        idx[pos] = idx[pos - 1];
                        --pos;
                    }
                    key[pos] = key_j;
                    idx[pos] = j;
                    if (filled < k)
                        ++filled;
                }
                for (int j = 0; j < k; ++j) {
                    knn_dist[(size_t)j * query_nb + i] = key[j];
                    knn_index[(size_t)j * query_nb + i] = idx[j];
                }
            }
        }
        free(key);
        free(idx);
    }
    return ok;
}

/*
 * The reference's acceptance check (code/test.cpp:216-217, :256-272): over the
 * k x query_nb grid, the fraction of positions whose distance is within
 * `precision` (absolute) of the ground truth and the fraction whose index is
 * equal.  The reference passes when both fractions are >= 0.999.
 */
void oracle_check(const float *test_dist, const int *test_index, const float *gt_dist,
                  const int *gt_index, long n, float precision, double *dist_frac,
                  double *index_frac) {
    long ok_d = 0, ok_i = 0;
    for (long i = 0; i < n; ++i) {
        if (fabs(test_dist[i] - gt_dist[i]) <= precision)
            ++ok_d;
        if (test_index[i] == gt_index[i])
            ++ok_i;
    }
    *dist_frac = n ? (double)ok_d / (double)n : 1.0;
    *index_frac = n ? (double)ok_i / (double)n : 1.0;
}

/* ---------------------------------------------------------------------------
 * Seeded generators.  splitmix64 of (seed, stream, counter) gives every value a
 * fixed position-independent source, so the HIP generator in the library and
 * this one agree for uniform data bit for bit.
 * ------------------------------------------------------------------------- */
static uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t oracle_rng_u64(uint64_t seed, uint64_t stream, uint64_t counter) {
    return mix64(seed * 0x9E3779B97F4A7C15ull + stream * 0xD1B54A32D192ED03ull +
                 counter * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull);
}

/* Uniform float in [0, 1): the top 24 bits of the counter hash. */
static float unit24(uint64_t u) { return (float)(u >> 40) * (1.0f / 16777216.0f); }

/* Point set of n points in `dim` dimensions, dim x n row-major: element (d, i)
 * at [d*n + i], value = scale * U[0,1) drawn from counter i*dim + d. */
void oracle_gen_uniform(float *out, int n, int dim, uint64_t seed, float scale) {
    for (int d = 0; d < dim; ++d)
        for (int i = 0; i < n; ++i)
            out[(size_t)d * n + i] =
                scale * unit24(oracle_rng_u64(seed, 0, (uint64_t)i * dim + d));
}

/* Gaussian mixture: `ncomp` isotropic components with centres uniform in
 * [-1, 1]^dim (drawn from `centre_seed`, so a reference set and a query set
 * that share it come from the same mixture), a uniform choice of component
 * per point and standard deviation `sigma`.  Normals come from Box-Muller in
 * double.  The HIP generator uses the same formula; its values may differ
 * from these in the last bit where the device log/cos round differently, so
 * tests always hand the oracle the data the device produced. */
void oracle_gen_mixture(float *out, int n, int dim, int ncomp, float sigma,
                        uint64_t centre_seed, uint64_t point_seed) {
    for (int i = 0; i < n; ++i) {
        uint64_t c = ((oracle_rng_u64(point_seed, 2, (uint64_t)i) >> 32) * (uint64_t)ncomp) >> 32;
        for (int d = 0; d < dim; ++d) {
            float centre = 2.0f * unit24(oracle_rng_u64(centre_seed, 1, c * (uint64_t)dim + d)) - 1.0f;
            uint64_t cnt = (uint64_t)i * dim + d;
            double u1 = ((double)(oracle_rng_u64(point_seed, 3, cnt) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
            double u2 = (double)(oracle_rng_u64(point_seed, 4, cnt) >> 11) * (1.0 / 9007199254740992.0);
            double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
            out[(size_t)d * n + i] = centre + sigma * (float)z;
        }
    }
}
