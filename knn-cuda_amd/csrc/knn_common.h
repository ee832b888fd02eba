// Shared device helpers and launch descriptors of the exact k-NN engine.
//
// Layout conventions (all device buffers):
//   point sets   dim-major: coordinate d of point i at p[d*ld + i]
//   results      k x Q: rank j of query i at out[j*Q + i]
//   partials     [slice][query][kp] candidate heaps written by the scan
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace knnhip {

constexpr int kWave = 64;            // CDNA wavefront width
constexpr int kQueriesPerWave = 32;  // one MFMA 32-column block of queries per wave
// One-wave ball tasks with merged candidate lists: queries a task, pending
// candidates a lane, candidate slot columns (both lane halves), and the
// largest kp (a merge's union, kp entries plus both halves' candidates,
// spans two passes of 64 lanes at most).
constexpr int kMergeQueries = 16;
constexpr int kMergeCand = 32;
constexpr int kMergeSlots = 2 * kMergeQueries;
constexpr int kMergeMaxKp = 2 * kWave - 2 * kMergeCand;
constexpr int kRefsPerTile = 32;     // one MFMA 32-row block of references
constexpr int kMTiles = 4;           // reference tiles per scan step (128 refs)
constexpr int kRefsPerStep = kRefsPerTile * kMTiles;
// Where the FP32 scan keeps the -2q operands: kBRegs, the padded dimension's
// values in VGPRs (one chunk of 2*KS dims); kBChunks, chunks of 2*KS dims
// reloaded into VGPRs chunk by chunk (any dimension); kBLds, the block's 32
// query rows staged in LDS once and read 16 bytes per dimension octet, every
// octet in one pass.
constexpr int kBRegs = 0, kBChunks = 1, kBLds = 2;
constexpr int kMaxUnion = 2048;      // max slices*kp entries merged per query
constexpr int kMaxCand = 1024;       // max candidates re-ranked exactly per query

// One scan task: references [ref_lo, ref_end) against up to 32 queries
// qlist[qoff .. qoff+qcount).  ref_begin <= ref_lo is ref_lo rounded down to
// a multiple of 32 (the packed tiles start there); rows in [ref_begin,
// ref_lo) take no part.  unit: the block (or ball) the range belongs to.
struct ScanTask {
    int ref_begin;
    int ref_end;
    int qoff;
    int qcount;
    int ref_lo;
    int unit;
};

// Arguments of the FP32 MFMA scan (one launch covers all queries and slices).
struct ScanArgs {
    const float *packed;  // packed reference tiles (see knn_scan.hip)
    int oct;              // dimension octets per point in the packed layout
    int R;
    int dim;
    const float *query;   // dim rows of ldq floats
    long long ldq;
    int Q;
    const float *rnorm;   // R squared norms (FP32)
    unsigned *theta;      // Q shared thresholds (order-mapped float codes)
    int slice_len;        // references per slice, multiple of kRefsPerStep
    long long slice_span; // packed bytes of one slice (< 4 GiB)
    int r_pad;            // references covered by the packed layout
    int nslices;
    int kp;               // heap capacity per (query, slice)
    int nqb;              // query blocks (waves_per_block * 32 queries each)
    int nchunks;          // dimension chunks of 2*KS dims
    int xcd_map;          // 1: keep each slice on one XCD (nslices % 8 == 0)
    float *part_val;      // [row][kp]; plain searches: row = slice * nqb * 32 + query
    int *part_idx;
    const ScanTask *tasks;  // explicit tasks (nullptr: the plain slice x query-block grid)
    int ntasks;           // one block per task
    const int *qlist;     // query ids referenced by the tasks
    int slice0;           // plain searches: first slice of this launch (a chunked
                          // upload scans each chunk's slices as it lands)
    const unsigned short *packed_bf;  // bfloat16 hi/lo tiles (precision scan), else unused
    // Point-major query rows (row stride query_pm_ld, a multiple of 4, pads
    // zero), or nullptr.  Task-mode searches pass them: a lane then reads its
    // query's operands as 16-byte pieces of one row instead of scattered
    // columns of the dim-major matrix.
    const float *query_pm;
    int query_pm_ld;
    // Residual form (block-index tasks): bfloat16 residual tiles r - c of
    // the block-ordered positions, their FP32 squared norms, the block
    // centroids (rows of query_pm_ld floats) and radii, the block of each
    // 32-position tile, and per heap row (task_row + column) four floats:
    // the row's distance bound from the triangle inequality, eps and eta
    // (see knn_scan_resid.hip).
    const unsigned short *packed_res;
    const float *res_norm;
    const float *res_cent;
    const float *res_rho;
    float *res_row;
    float res_eta_coef;  // key error per (||e_q|| + ||e_r||)^2
    float res_norm_up;   // exact norm over its FP32 sum, at most
    float res_norm_dn;   // ... at least
};

// Arguments of the norms / layout pass.
struct NormArgs {
    const float *p;      // dim rows of ld floats
    long long ld;
    int n;
    int n_pad;           // points covered (packed tail tiles are zeroed)
    int dim;
    int oct;             // dimension octets written
    float *out;          // n squared norms
    unsigned *maxbits;   // max norm bits, or nullptr
    float *pm;           // point-major copy (row stride pm_ld), or nullptr
    int pm_ld;
    float *packed;       // packed tile layout, or nullptr
    unsigned short *packed_bf;  // bfloat16 hi/lo tile layout of the precision scan, or nullptr
    int nch_bf;                 // its 16-dimension chunks per tile (>= ceil(dim / 16))
    unsigned *nonfinite;        // set to 1 when a point's squared norm is NaN or infinite, or nullptr
};

// One point set of a multi-set norms pass: its points land at `start` (a
// multiple of 32) in the pass's outputs (norms, point-major rows, tiles).
struct NormJob {
    const float *p;  // dim rows of ld floats
    long long ld;
    int n;
    int n_pad;       // points covered (a multiple of 32 for tiles)
    int start;
};

// Arguments of the select / exact re-rank / certificate stage.
struct FinalArgs {
    const float *ref;
    long long ldr;
    int R;
    const float *query;
    long long ldq;
    int Q;
    int dim;
    int k;
    int kp;
    int nrows;                // heap rows per query
    long long row_stride;     // row of (query q, j) = j * row_stride + q when rows == nullptr
    const int *rows;          // else rows[q * nrows + j] (negative: unused)
    const int *perm;          // reference position -> caller's index (nullptr: identity)
    int maxc;                 // candidate capacity per query (<= kMaxCand)
    const float *part_val;
    const int *part_idx;
    const float *ref_pm;      // point-major copies: point i at [i*pm_ld], dim floats
    const float *query_pm;
    int pm_ld;                // multiple of 4 (16-byte rows)
    const float *qnorm;       // Q squared norms (FP32)
    const unsigned *rmax;     // bit pattern of max_r ||r||^2 (non-negative float)
    const unsigned *theta;    // final shared thresholds of the scan (order-mapped)
    const float *lb_rest;     // per query, a lower bound on the distance (not squared) of
                              // every reference the scan skipped (nullptr: none skipped)
    int force_fallback;
    float *out_dist;          // k x Q
    int *out_idx;
    int *fb_count;            // queries that failed the certificate
    int *fb_list;
    // Key error bound of the scan: |key - (||r||^2 - 2 r.q)| <= err_q * ||q||^2 +
    // err_r * max ||r||^2.  Zero: the FP32 scan's bound (g and 3.1 g).
    double err_q, err_r;
    // Non-zero when an input point has a NaN or infinite squared norm: every
    // output of the search is then index -1, distance NaN (nullptr: unchecked).
    const unsigned *nonfinite;
    // Residual-form scan (non-null): keys are approximate squared distances;
    // per heap row four floats (L, eps, eta, -): every reference of the
    // row's block is at distance >= L, and every one the scan rejected is
    // also at distance >= sqrt(tau - eta) - eps.
    const float *res_row;
};

// Arguments of the exact per-query full scan (fallback).
struct ExactArgs {
    const float *ref;
    long long ldr;
    int R;
    const float *query;
    long long ldq;
    int Q;
    int dim;
    int k;
    const int *list;     // query ids to answer (nullptr: all queries 0..Q-1)
    const int *count;    // number of ids in list (nullptr: Q)
    float *scratch_key;  // [gridDim.x][64][k]
    int *scratch_idx;
    float *out_dist;
    int *out_idx;
    const unsigned *nonfinite;  // as FinalArgs::nonfinite
};

// The block-index searches' bound matrix: bound (b, q) of block b for query
// q, in slabs of 64 queries (slab q / 64 holds nb x 64 bounds, block-major),
// so the 64 queries of a pick workgroup read one contiguous slab.
__host__ __device__ inline long long lb_at(int b, int q, int nb) {
    return ((long long)(q >> 6) * nb + b) * 64 + (q & 63);
}
__host__ __device__ inline long long lb_elems(int Q, int nb) { return (long long)((Q + 63) / 64) * 64 * nb; }

// Four floats of a point-major row at d4 (a multiple of 4), zeros at and
// past ld (the row stride: a multiple of 4, at least 4).  The load itself is
// unconditional (clamped to the row's last piece), so a run of these issues
// back to back instead of one branch and wait per piece.
__device__ __forceinline__ float4 row_piece(const float *row, int d4, int ld) {
    const float4 v = *reinterpret_cast<const float4 *>(row + min(d4, ld - 4));
    return d4 < ld ? v : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}
// Order-preserving map of a float onto uint32 (total order, -0 < +0).
__device__ __forceinline__ unsigned float_to_ordered(float x) {
    unsigned b = __float_as_uint(x);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ordered_to_float(unsigned u) {
    unsigned b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    return __uint_as_float(b);
}

// Raw buffer resource over [base, base + nbytes) from values that are
// uniform over the wave.  Buffer loads take their resource in scalar
// registers; values the compiler holds in vector registers (after a branch
// merge, or read from LDS) would make it wrap every load in a loop over the
// distinct resources of the lanes, so the base and size are read from the
// first lane here.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t uniform_buffer_rsrc(const void *base, unsigned nbytes) {
    const unsigned long long p = (unsigned long long)base;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)p);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(p >> 32));
    return __builtin_amdgcn_make_buffer_rsrc((void *)(((unsigned long long)hi << 32) | lo), (short)0,
                                             (int)__builtin_amdgcn_readfirstlane(nbytes), 0x00020000);
}

// Compiler-level ordering point for LDS traffic shared between lanes of one
// wave (no hardware barrier: waves of a block never wait for each other).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Squared distance exactly as the reference ground truth forms it
// (code/test.cpp:61-66): ascending d, FP32 difference, FP32 square, FP32 sum,
// no contraction into FMA.
__device__ __forceinline__ float canonical_sq(const float *__restrict__ ref, long long ldr,
                                              int r, const float *__restrict__ query,
                                              long long ldq, int q, int dim) {
#pragma clang fp contract(off)
    float acc = 0.0f;
    const float *rp = ref + r;
    const float *qp = query + q;
    int d = 0;
    for (; d + 8 <= dim; d += 8) {
        float rv[8], qv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            rv[j] = rp[(long long)(d + j) * ldr];
            qv[j] = qp[(long long)(d + j) * ldq];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            float t = rv[j] - qv[j];
            float s = t * t;
            acc = acc + s;
        }
    }
    for (; d < dim; ++d) {
        float t = rp[(long long)d * ldr] - qp[(long long)d * ldq];
        float s = t * t;
        acc = acc + s;
    }
    return acc;
}

// The same ground-truth sum over two contiguous point-major rows.
__device__ __forceinline__ float canonical_sq_rows(const float *__restrict__ r,
                                                   const float *__restrict__ q, int dim) {
#pragma clang fp contract(off)
    float acc = 0.0f;
    int d = 0;
    for (; d + 4 <= dim; d += 4) {
        const float4 rv = *reinterpret_cast<const float4 *>(r + d);
        const float4 qv = *reinterpret_cast<const float4 *>(q + d);
        float t = rv.x - qv.x;
        float s = t * t;
        acc = acc + s;
        t = rv.y - qv.y;
        s = t * t;
        acc = acc + s;
        t = rv.z - qv.z;
        s = t * t;
        acc = acc + s;
        t = rv.w - qv.w;
        s = t * t;
        acc = acc + s;
    }
    for (; d < dim; ++d) {
        const float t = r[d] - q[d];
        const float s = t * t;
        acc = acc + s;
    }
    return acc;
}

// Correctly rounded FP32 square root: the double square root of a float
// argument rounded once to float equals the correctly rounded float result.
__device__ __forceinline__ float sqrt_rn(float x) { return (float)__builtin_sqrt((double)x); }

// Lexicographic (key, index) order of the reference top-k.
__device__ __forceinline__ bool lex_less(float ka, int ia, float kb, int ib) {
    return ka < kb || (ka == kb && ia < ib);
}

}  // namespace knnhip
