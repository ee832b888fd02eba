// Transfers between the caller's (pageable) host arrays and the device that
// are ordered on a stream like any kernel.
//
// The calling thread gathers host rows into one of two pinned staging
// buffers (hipHostMalloc); a linear DMA moves the piece into a device
// staging buffer and a kernel writes it into the strided destination (or the
// DMA goes straight to a contiguous destination).  Asynchronous copies from
// pageable memory are not used: their completion was seen not to be covered
// by events recorded behind them on the same stream.
//
// host_upload returns once every piece is queued; the caller's array is no
// longer read by then (it has been copied to pinned memory).  host_download
// returns when the destination holds the data.
#include <algorithm>
#include <thread>
#include <vector>

#include "knn_engine.h"

namespace knnhip {

namespace {

constexpr size_t kStageBytes = 32u << 20;  // per staging slot
constexpr int kCopyThreads = 8;
constexpr size_t kBytesPerThread = 4u << 20;

struct Staging {
    char *host[2] = {};
    DevBuf dev[2];
    hipEvent_t done[2] = {};
    int next = 0;
    bool ready = false;
};
Staging g_st;

bool stage_init() {
    if (g_st.ready) return true;
    for (int i = 0; i < 2; ++i) {
        KNN_TRY(hipHostMalloc((void **)&g_st.host[i], kStageBytes, hipHostMallocDefault), "host alloc");
        KNN_TRY(hipEventCreateWithFlags(&g_st.done[i], hipEventDisableTiming), "event create");
        KNN_TRY(hipEventRecord(g_st.done[i], nullptr), "event");
    }
    g_st.ready = true;
    return true;
}

// Device staging of a slot, for strided device layouts only.
bool stage_dev(int slot) {
    if (g_st.dev[slot].p) return true;
    if (!g_st.dev[slot].alloc(kStageBytes)) {
        set_error("Memory allocation error");
        return false;
    }
    return true;
}

// f(i) for i in [0, n), split over up to kCopyThreads threads when the piece
// is large enough to pay for them.
template <class F> void parallel_for(int n, size_t bytes, F f) {
    const int t = (int)std::min<size_t>({(size_t)kCopyThreads, (size_t)n, bytes / kBytesPerThread + 1});
    if (t <= 1) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    // a thread that cannot be started leaves its share to this one (no
    // exception crosses the C interface)
    std::vector<std::thread> th;
    try {
        th.reserve(t);
    } catch (...) {
    }
    for (int j = 0; j < t; ++j) {
        try {
            th.emplace_back([=]() {
                for (int i = (int)((long long)n * j / t); i < (int)((long long)n * (j + 1) / t); ++i) f(i);
            });
        } catch (...) {
            for (int i = (int)((long long)n * j / t); i < (int)((long long)n * (j + 1) / t); ++i) f(i);
        }
    }
    for (auto &x : th) x.join();
}

// Rows of a pitched array (4-byte words) to or from a packed piece of nr
// rows of w words: blocks over (columns, rows), no index division.
constexpr int kMoveThreads = 256;
__global__ __launch_bounds__(kMoveThreads) void scatter_rows_kernel(const unsigned *__restrict__ src,
                                                                    unsigned *__restrict__ dst,
                                                                    long long dld, long long w, int nr) {
    for (int r = blockIdx.y; r < nr; r += gridDim.y)
        for (long long c = (long long)blockIdx.x * kMoveThreads + threadIdx.x; c < w;
             c += (long long)gridDim.x * kMoveThreads)
            dst[(long long)r * dld + c] = src[(long long)r * w + c];
}
__global__ __launch_bounds__(kMoveThreads) void gather_rows_kernel(const unsigned *__restrict__ src,
                                                                   long long sld, long long w, int nr,
                                                                   unsigned *__restrict__ dst) {
    for (int r = blockIdx.y; r < nr; r += gridDim.y)
        for (long long c = (long long)blockIdx.x * kMoveThreads + threadIdx.x; c < w;
             c += (long long)gridDim.x * kMoveThreads)
            dst[(long long)r * w + c] = src[(long long)r * sld + c];
}
dim3 move_grid(long long w, int nr) {
    const long long gx = std::min<long long>((w + kMoveThreads - 1) / kMoveThreads, 1024);
    const long long gy = std::min<long long>(nr, std::max<long long>(1, 4096 / gx));
    return dim3((unsigned)std::max(1ll, gx), (unsigned)std::max(1ll, gy));
}

// Pieces of at most kStageBytes: whole rows when a row fits, else segments
// of one row.
template <class F> bool for_pieces(long long width, int rows, F f) {
    const long long slot_words = (long long)(kStageBytes / 4);
    if (width <= slot_words) {
        const int per = (int)std::min<long long>(rows, slot_words / width);
        for (int r = 0; r < rows; r += per)
            if (!f(r, std::min(per, rows - r), 0LL, width)) return false;
    } else {
        for (int r = 0; r < rows; ++r)
            for (long long c = 0; c < width; c += slot_words)
                if (!f(r, 1, c, std::min(slot_words, width - c))) return false;
    }
    return true;
}

}  // namespace

bool host_upload(void *dst, long long dld, const void *src, long long sld, long long width, int rows,
                 hipStream_t s) {
    if (width <= 0 || rows <= 0) return true;
    if (!stage_init()) return false;
    unsigned *d = (unsigned *)dst;
    const unsigned *h = (const unsigned *)src;
    return for_pieces(width, rows, [&](int r0, int nr, long long c0, long long w) -> bool {
        const int slot = g_st.next;
        g_st.next ^= 1;
        // the slot's previous DMA and scatter are done with it
        KNN_TRY(hipEventSynchronize(g_st.done[slot]), "event sync");
        unsigned *stage = (unsigned *)g_st.host[slot];
        const size_t bytes = (size_t)nr * w * 4;
        if (nr == 1) {
            const int t = (int)std::min<size_t>(kCopyThreads, bytes / kBytesPerThread + 1);
            parallel_for(t, bytes, [&](int j) {
                const long long a = w * j / t, b = w * (j + 1) / t;
                memcpy(stage + a, h + (size_t)r0 * sld + c0 + a, (size_t)(b - a) * 4);
            });
        } else {
            parallel_for(nr, bytes, [&](int i) {
                memcpy(stage + (size_t)i * w, h + (size_t)(r0 + i) * sld + c0, (size_t)w * 4);
            });
        }
        unsigned *out = d + (size_t)r0 * dld + c0;
        if (nr == 1 || dld == w) {
            KNN_TRY(hipMemcpyAsync(out, stage, bytes, hipMemcpyHostToDevice, s),
                    "Unable to copy data from host to device");
        } else {
            if (!stage_dev(slot)) return false;
            KNN_TRY(hipMemcpyAsync(g_st.dev[slot].p, stage, bytes, hipMemcpyHostToDevice, s),
                    "Unable to copy data from host to device");
            hipLaunchKernelGGL(scatter_rows_kernel, move_grid(w, nr), dim3(kMoveThreads), 0, s,
                               g_st.dev[slot].as<unsigned>(), out, dld, w, nr);
            KNN_TRY(hipGetLastError(), "copy kernel launch");
        }
        KNN_TRY(hipEventRecord(g_st.done[slot], s), "event");
        return true;
    });
}

bool host_download(void *dst, long long dld, const void *src, long long sld, long long width, int rows,
                   hipStream_t s) {
    if (width <= 0 || rows <= 0) return true;
    if (!stage_init()) return false;
    unsigned *h = (unsigned *)dst;
    const unsigned *d = (const unsigned *)src;
    return for_pieces(width, rows, [&](int r0, int nr, long long c0, long long w) -> bool {
        const int slot = g_st.next;
        g_st.next ^= 1;
        KNN_TRY(hipEventSynchronize(g_st.done[slot]), "event sync");
        unsigned *stage = (unsigned *)g_st.host[slot];
        const size_t bytes = (size_t)nr * w * 4;
        const unsigned *in = d + (size_t)r0 * sld + c0;
        if (nr == 1 || sld == w) {
            KNN_TRY(hipMemcpyAsync(stage, in, bytes, hipMemcpyDeviceToHost, s),
                    "Unable to copy data from device to host");
        } else {
            if (!stage_dev(slot)) return false;
            hipLaunchKernelGGL(gather_rows_kernel, move_grid(w, nr), dim3(kMoveThreads), 0, s, in, sld, w, nr,
                               g_st.dev[slot].as<unsigned>());
            KNN_TRY(hipGetLastError(), "copy kernel launch");
            KNN_TRY(hipMemcpyAsync(stage, g_st.dev[slot].p, bytes, hipMemcpyDeviceToHost, s),
                    "Unable to copy data from device to host");
        }
        KNN_TRY(hipEventRecord(g_st.done[slot], s), "event");
        KNN_TRY(hipEventSynchronize(g_st.done[slot]), "Unable to copy data from device to host");
        if (nr == 1) {
            const int t = (int)std::min<size_t>(kCopyThreads, bytes / kBytesPerThread + 1);
            parallel_for(t, bytes, [&](int j) {
                const long long a = w * j / t, b = w * (j + 1) / t;
                memcpy(h + (size_t)r0 * dld + c0 + a, stage + a, (size_t)(b - a) * 4);
            });
        } else {
            parallel_for(nr, bytes, [&](int i) {
                memcpy(h + (size_t)(r0 + i) * dld + c0, stage + (size_t)i * w, (size_t)w * 4);
            });
        }
        return true;
    });
}

}  // namespace knnhip
