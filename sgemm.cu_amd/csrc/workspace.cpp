// workspace.cpp -- split-K workspace per stream (with the BF16x3 packed
// operand planes) and the split count rule.
//
// Split-K partials and the per-tile arrival counters live in device memory
// owned by the library: one workspace per stream (launches on one stream are
// ordered, so they can share it; launches on different streams get their own).
// The counters are zeroed once when allocated (on the owning stream, ahead of
// its first launch) and every split-K launch leaves
// them zero again, so no memset is needed per launch (and none lands in a
// captured graph).  A stream that is being captured never allocates: a
// split-K launch whose workspace would have to grow runs unsplit instead, so
// warm a stream up before capturing on it to keep the split.  A workspace
// that is outgrown is retired, not freed: a captured graph may still hold its
// address, so it stays valid until the stream's workspace is released
// (splitk_workspace_release) or the process ends.  Growth is geometric, so
// the retired total stays below the live workspace.
//
// Launches that share a workspace must not overlap: eager calls on one
// stream are ordered by the stream, but a graph that captured split launches
// on that stream uses the same partials and counters when it is replayed, so
// its replays must be ordered with each other and with eager calls on the
// capture stream (a counter left non-zero by an overlapping launch makes a
// later launch skip writing tiles).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "sgemm.h"
#include "kernels.hpp"

namespace sgemm_amd {

namespace {

struct Workspace {
    void* partials = nullptr;
    size_t partial_bytes = 0;
    int* counters = nullptr;
    int counter_count = 0;
    void* packed = nullptr;      // BF16x3 config 12's packed operand planes
    size_t packed_bytes = 0;
    std::vector<void*> retired;  // outgrown buffers, kept for captured graphs
};

// True when `stream` is being captured into a graph.
bool capturing(hipStream_t stream) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &cs) != hipSuccess) (void)hipGetLastError();
    return cs != hipStreamCaptureStatusNone;
}

std::mutex g_mu;
std::map<hipStream_t, Workspace> g_ws;

}  // namespace

int device_cu_count() {
    static int cus = [] {
        int dev = 0, n = 256;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
            return n;
        (void)hipGetLastError();
        return 256;
    }();
    return cus;
}

int splitk_count(long long tiles, int k, int bk, bool f32, int min_ktiles) {
    if (f32) {
        // Measured on the square sizes 1024..3072 (config 7): 1152^3 63 -> 91,
        // 1280^3 78 -> 90, 1536^3 84 -> 97 TFLOP/s over the two-per-CU rule,
        // and no loss on 512 log-uniform problems of 32..2048 launched one by
        // one (a cap of eight k-tiles per split cost those 17 %).
        const double want = 4.0 * device_cu_count() / (double)std::max(tiles, 1LL);
        const long long s = std::min({(long long)(want + 0.5),
                                      (long long)std::max(1, k / (min_ktiles * bk)), 16LL});
        return (int)std::max(1LL, s);
    }
    const long long target = 2LL * device_cu_count();
    if (tiles >= target) return 1;
    long long s = target / std::max(tiles, 1LL);
    const long long max_by_k = std::max(1, k / (min_ktiles * bk));
    s = std::min({s, max_by_k, 16LL});
    return (int)std::max(1LL, s);
}

int tail_split_count(long long tiles, long long slots, int per_cu, int k, int bk,
                     size_t split_bytes, double round_us, double* cost) {
    const long long cus = std::max(1LL, slots / std::max(1, per_cu));
    const long long rem = slots > 0 ? tiles % slots : 0;
    *cost = (double)((rem + cus - 1) / cus) / std::max(1, per_cu);
    if (tiles < slots || rem == 0 || round_us <= 0) return 1;
    const int max_s = std::min(16, k / (4 * bk));
    int best = 1;
    for (int s = 2; s <= max_s; ++s) {
        if ((double)rem * s * (double)split_bytes > 512.0 * 1024 * 1024) break;
        const long long rounds = (rem * s + slots - 1) / slots;
        const double traffic_us = (double)(s + 1) * (double)split_bytes / 32.0e3;
        const double c = (double)rounds / s + traffic_us / round_us;
        if (c < 0.97 * *cost) {
            *cost = c;
            best = s;
        }
    }
    return best;
}

long long streamk_iters(long long rem, long long slots, long long iters, size_t split_bytes,
                        double piece_iters, double* part) {
    *part = 1.0;
    if (rem <= 0 || slots <= 0 || iters <= 0) return 0;
    const long long L = rem * iters;
    const long long w0 = std::max(1LL, (L + slots - 1) / slots);
    // For a piece bound p (2..16) the cheapest range length is the shortest
    // w >= w0 with at most p pieces per tile, (iters - 1) / w + 2 <= p, i.e.
    // w >= (iters - 1) / (p - 1) + 1, and with the partials of ceil(L / w)
    // workgroups x p pieces within the cap; w0 itself is the other candidate.
    constexpr double kCap = 512.0 * 1024 * 1024;
    long long best = 0;
    double best_t = 0;
    for (int p = 1; p <= 16; ++p) {
        long long w = w0;
        if (p > 1) {
            const long long nmax = (long long)(kCap / ((double)p * split_bytes));
            if (nmax < 1) continue;
            w = std::max({w0, (iters - 1) / (p - 1) + 1, (L + nmax - 1) / nmax});
        }
        const long long pieces = (iters - 1) / w + 2;
        if (pieces > 16) continue;
        if ((double)((L + w - 1) / w) * pieces * split_bytes > kCap) continue;
        const double t = (double)w + piece_iters * pieces;
        if (best == 0 || t < best_t || (t == best_t && w < best)) {
            best = w;
            best_t = t;
        }
    }
    if (best > 0) *part = best_t / iters;
    return best;
}

hipError_t streamk_plan(hipStream_t stream, long long tiles, long long slots, int k, int bk,
                        size_t split_bytes, double piece_iters, SplitK* sk, long long* nwg,
                        const Arena* arena) {
    *nwg = 0;
    const long long rem = slots > 0 ? tiles % slots : 0;
    if (rem <= 0 || k <= 0) return hipSuccess;
    const long long iters = div_up(k, bk);
    double part = 1.0;
    const long long w = streamk_iters(rem, slots, iters, split_bytes, piece_iters, &part);
    if (w <= 0) return hipSuccess;
    const long long n = (rem * iters + w - 1) / w;
    if (n <= 1) return hipSuccess;
    sk->iters = (int)w;
    sk->pieces = (int)((iters - 1) / w + 2);
    const hipError_t e = splitk_workspace(stream, (size_t)n * sk->pieces * split_bytes, (int)n,
                                          &sk->partials, &sk->counters, arena);
    if (e == hipErrorStreamCaptureUnsupported) return hipSuccess;
    if (e != hipSuccess) return e;
    *nwg = n;
    return hipSuccess;
}

hipError_t splitk_plan(hipStream_t stream, long long tiles, int k, int bk, size_t split_bytes,
                       SplitK* sk, bool f32, int min_ktiles, const Arena* arena) {
    sk->partials = nullptr;
    sk->counters = nullptr;
    sk->splits = splitk_count(tiles, k, bk, f32, min_ktiles);
    sk->kchunk = div_up(div_up(k, sk->splits), bk) * bk;
    sk->splits = div_up(k, sk->kchunk);  // every split gets some k
    if (sk->splits <= 1) return hipSuccess;
    const hipError_t e = splitk_workspace(stream, (size_t)tiles * sk->splits * split_bytes,
                                          (int)tiles, &sk->partials, &sk->counters, arena);
    if (e == hipErrorStreamCaptureUnsupported) {
        // Capturing, and this stream's workspace is too small: run unsplit
        // rather than allocate inside the capture.
        sk->splits = 1;
        sk->kchunk = k;
        sk->partials = nullptr;
        sk->counters = nullptr;
        return hipSuccess;
    }
    return e;
}

hipError_t splitk_workspace(hipStream_t stream, size_t partial_bytes, int tiles,
                            float** partials, int** counters, const Arena* arena) {
    if (arena != nullptr) {
        if (partial_bytes > arena->partial_bytes || tiles > arena->counter_count)
            return hipErrorOutOfMemory;
        *partials = arena->partials;
        *counters = arena->counters;
        return hipSuccess;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    Workspace& w = g_ws[stream];
    if ((w.partial_bytes < partial_bytes || w.counter_count < tiles) && capturing(stream))
        return hipErrorStreamCaptureUnsupported;
    if (w.partial_bytes < partial_bytes) {
        const size_t bytes = std::max(partial_bytes + partial_bytes / 4, 2 * w.partial_bytes);
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return hipErrorOutOfMemory;
        }
        if (w.partials) w.retired.push_back(w.partials);
        w.partials = p;
        w.partial_bytes = bytes;
    }
    if (w.counter_count < tiles) {
        const int n = std::max({tiles, 4096, 2 * w.counter_count});
        int* c = nullptr;
        if (hipMalloc(&c, (size_t)n * sizeof(int)) != hipSuccess) {
            (void)hipGetLastError();
            return hipErrorOutOfMemory;
        }
        // Zeroed on the stream that will use the counters: a plain hipMemset
        // runs on the null stream, which a non-blocking stream does not wait
        // for, so a launch right after it could read recycled memory as
        // arrival counts (no workgroup would see itself last, and the tiles
        // would never be written).  Never reached under capture.
        if (hipMemsetAsync(c, 0, (size_t)n * sizeof(int), stream) != hipSuccess) {
            (void)hipFree(c);
            return hipErrorUnknown;
        }
        if (w.counters) w.retired.push_back(w.counters);
        w.counters = c;
        w.counter_count = n;
    }
    *partials = static_cast<float*>(w.partials);
    *counters = w.counters;
    return hipSuccess;
}

hipError_t pack_workspace(hipStream_t stream, size_t bytes, void** out) {
    std::lock_guard<std::mutex> lock(g_mu);
    Workspace& w = g_ws[stream];
    if (w.packed_bytes < bytes) {
        if (capturing(stream)) return hipErrorStreamCaptureUnsupported;
        const size_t grow = std::max(bytes + bytes / 4, 2 * w.packed_bytes);
        void* p = nullptr;
        if (hipMalloc(&p, grow) != hipSuccess) {
            (void)hipGetLastError();
            return hipErrorOutOfMemory;
        }
        if (w.packed) w.retired.push_back(w.packed);
        w.packed = p;
        w.packed_bytes = grow;
    }
    *out = w.packed;
    return hipSuccess;
}

hipError_t splitk_workspace_release(hipStream_t stream) {
    Workspace w;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        auto it = g_ws.find(stream);
        if (it == g_ws.end()) return hipSuccess;
        w = std::move(it->second);
        g_ws.erase(it);
    }
    // The stream's pending launches may still use the buffers.
    hipError_t e = hipStreamSynchronize(stream);
    if (w.partials) (void)hipFree(w.partials);
    if (w.counters) (void)hipFree(w.counters);
    if (w.packed) (void)hipFree(w.packed);
    for (void* p : w.retired) (void)hipFree(p);
    return e;
}

}  // namespace sgemm_amd

extern "C" int sgemm_workspace_release(sgemm_stream_t stream) {
    const hipError_t e = sgemm_amd::splitk_workspace_release(static_cast<hipStream_t>(stream));
    if (e == hipSuccess) return SGEMM_STATUS_SUCCESS;
    (void)hipGetLastError();
    return SGEMM_STATUS_HIP_ERROR;
}
