// grouped.hip -- one persistent launch for many independent GEMM problems of
// uneven shape, plus the plan object that schedules them and captures the
// launch in a hipGraph.  No reference counterpart (the reference launches one
// problem per call, src/sgemm.cuh:40-47).
//
// Schedule.  Every problem is cut into tiles of the config the autotune table
// picks for its (form, m, n, k) size bucket.  Problems are ordered by the cost
// of one of their tiles (k * bm * bn, largest first; ties keep submission
// order), and their tiles are concatenated into one index space.  The
// persistent workgroups take tile indices from a device-wide queue (an
// agent-scope atomic counter in the header), so a workgroup that finishes a
// cheap tile takes the next one at once: longest-processing-time-first list
// scheduling, which leaves at most one of the cheapest tiles of imbalance at
// the end.  (A static stride b, b + G, b + 2G, ... gives workgroup 0 the most
// expensive tile of every pass; on 512 problems of log-uniform size that
// imbalance cost 10-20 %.)  The last workgroup to leave resets the counters,
// so every launch -- and every replay of a captured graph -- starts from zero
// without a memset.
//
// Each tile is computed by the same tile code as the single-problem kernels
// (tile_f32.hpp / tile_x3.hpp) with the same config, so a problem's result is
// bit-identical whether it runs alone through sgemm_ex / sgemm_bf16x3 with
// that config or inside any group.
//
// The descriptor array lives in device memory behind a small header
// {count, total_tiles}; the kernel reads both from there, so a captured graph
// stays valid when the plan is re-filled with other problems (up to its
// capacity) -- only the descriptors change, not the launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "sgemm.h"
#include "autotune.hpp"
#include "grouped.hpp"
#include "tile_f32.hpp"
#include "tile_x3.hpp"

namespace sgemm_amd {

struct GroupHeader {
    int count;
    int total_tiles;
    int pad0[6];
    // Tile queue, zero between launches: the next tile index to hand out, and
    // the number of workgroups that have left.  Only atomics touch them.
    int next_tile;
    int exited;
    int pad1[6];  // keeps the descriptor array 64-byte aligned
};
static_assert(sizeof(GroupHeader) == 64, "header size");

// Configs a grouped kernel carries (grouped_config, all 256 threads).  The
// kernel's registers and LDS are the maximum over them: for FP32 that is
// 128x64x32's 52 KB of LDS and 128x128x16's registers, three workgroups per
// CU -- the 128x128x32 tile (70 KB, two per CU) is left out; for BF16x3 the
// single-accumulator 128x128 tile's 80 KB, two per CU.
constexpr bool grouped_has(int form, int cfg) { return grouped_config(form, cfg); }

constexpr int kGroupedLdsF32 = F32_128x64::LDS_FLOATS * 4;
constexpr int kGroupedLdsX3 = X3_128x128_2cu::LDS_BYTES;
static_assert(F32_64x64::LDS_FLOATS * 4 <= kGroupedLdsF32 &&
                  F32_128x128_k16::LDS_FLOATS * 4 <= kGroupedLdsF32,
              "F32 grouped LDS");
static_assert(X3_64x64::LDS_BYTES <= kGroupedLdsX3, "X3 grouped LDS");
constexpr int kGroupedWpeF32 = 3;
constexpr int kGroupedWpeX3 = 2;

template <class T>
__device__ inline void run_tile_f32(const Problem& P, int tm, int tn, char* smem) {
    F32TileRunner<T, false>::run(P, tm, tn, reinterpret_cast<float*>(smem));
}

template <int FORM>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(FORM == kFormF32 ? kGroupedWpeF32 : kGroupedWpeX3)))
void k_grouped(GroupHeader* __restrict__ hdr, const Problem* __restrict__ probs) {
    __shared__ __attribute__((aligned(16))) char smem[FORM == kFormF32 ? kGroupedLdsF32
                                                                        : kGroupedLdsX3];
    // The tile index is broadcast through the first word of the tile's own
    // LDS between tiles: a separate __shared__ word would push the BF16x3
    // kernel past the 80 KB that two workgroups per CU allow.
    int* const s_tile = reinterpret_cast<int*>(smem);
    const int count = hdr->count;
    const int total = hdr->total_tiles;
    const bool leader = threadIdx.x == 0;
    if (leader)
        *s_tile = __hip_atomic_fetch_add(&hdr->next_tile, 1, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    int w = *s_tile;
    while (w < total) {
        // Take the following index now; its latency hides behind this tile.
        int next = 0;
        if (leader)
            next = __hip_atomic_fetch_add(&hdr->next_tile, 1, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();  // every thread has read the index before the tile reuses the LDS
        int lo = 0, hi = count - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (probs[mid].tile_begin <= w) lo = mid;
            else hi = mid - 1;
        }
        const Problem P = probs[lo];
        int tm, tn;
        tile_coords(w - P.tile_begin, P.tiles_m, P.tiles_n, 8, tm, tn);
        if constexpr (FORM == kFormF32) {
            switch (P.cfg) {
                case 5: run_tile_f32<F32_128x128_k16>(P, tm, tn, smem); break;
                case 1: run_tile_f32<F32_64x64>(P, tm, tn, smem); break;
                default: run_tile_f32<F32_128x64>(P, tm, tn, smem); break;
            }
        } else {
            switch (P.cfg) {
                case 8: X3TileRunner<X3_128x128_2cu, false>::run(P, tm, tn, smem); break;
                default: X3TileRunner<X3_64x64, false>::run(P, tm, tn, smem); break;
            }
        }
        __syncthreads();  // the tile is done with the LDS
        if (leader) *s_tile = next;
        __syncthreads();
        w = *s_tile;
    }
    // Every index fetch of this workgroup precedes its exit count, so the
    // last workgroup out can clear the queue for the next launch.
    if (leader) {
        const int gone = __hip_atomic_fetch_add(&hdr->exited, 1, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
        if (gone == (int)gridDim.x - 1) {
            __hip_atomic_store(&hdr->next_tile, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&hdr->exited, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ------------------------------------------------------------------ host side
namespace {

int persistent_grid(int form) {
    static std::once_flag once;
    static int grid[2] = {0, 0};
    std::call_once(once, [] {
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            cus = 256;
        int nb = 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_grouped<kFormF32>, 256, 0) !=
                hipSuccess || nb < 1)
            nb = 1;
        grid[0] = cus * nb;
        nb = 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_grouped<kFormBf16x3>, 256, 0) !=
                hipSuccess || nb < 1)
            nb = 1;
        grid[1] = cus * nb;
        (void)hipGetLastError();
    });
    return grid[form == kFormF32 ? 0 : 1];
}

}  // namespace

int validate_problem(const sgemm_problem& q) {
    if (q.m < 0 || q.n < 0 || q.k < 0) return SGEMM_STATUS_INVALID_VALUE;
    if (q.lda < std::max(1, q.m) || q.ldb < std::max(1, q.k) || q.ldc < std::max(1, q.m))
        return SGEMM_STATUS_INVALID_VALUE;
    if (q.m > 0 && q.n > 0) {
        if (q.lda > kMaxLd || q.ldb > kMaxLd) return SGEMM_STATUS_NOT_SUPPORTED;
        if (q.C == nullptr || !dword_aligned(q.C)) return SGEMM_STATUS_INVALID_VALUE;
        if (q.k > 0 && q.alpha != 0.f &&
            (q.A == nullptr || q.B == nullptr || !dword_aligned(q.A) || !dword_aligned(q.B)))
            return SGEMM_STATUS_INVALID_VALUE;
    }
    return SGEMM_STATUS_SUCCESS;
}

Problem to_problem(const sgemm_problem& q) {
    Problem p;
    std::memset(&p, 0, sizeof(p));
    p.A = q.A;
    p.B = q.B;
    p.C = q.C;
    p.m = q.m;
    p.n = q.n;
    // alpha == 0 or k == 0: no product term, A and B are not read.
    p.k = (q.alpha == 0.f) ? 0 : q.k;
    p.lda = q.lda;
    p.ldb = q.ldb;
    p.ldc = q.ldc;
    p.alpha = q.alpha;
    p.beta = q.beta;
    return p;
}

int GroupedPlan::create(GroupedPlan** out, int capacity, int form) {
    if (out == nullptr || capacity < 1) return SGEMM_STATUS_INVALID_VALUE;
    if (form != kFormF32 && form != kFormBf16x3) return SGEMM_STATUS_NOT_SUPPORTED;
    GroupedPlan* p = new GroupedPlan();
    p->form_ = form;
    p->capacity_ = capacity;
    const size_t bytes = sizeof(GroupHeader) + sizeof(Problem) * (size_t)capacity;
    p->bytes_ = bytes;
    if (hipMalloc(&p->dev_, bytes) != hipSuccess ||
        hipHostMalloc(&p->staging_, bytes, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&p->uploaded_, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        p->release();
        delete p;
        return SGEMM_STATUS_ALLOC_FAILED;
    }
    std::memset(p->staging_, 0, bytes);
    if (hipMemcpy(p->dev_, p->staging_, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        p->release();
        delete p;
        return SGEMM_STATUS_HIP_ERROR;
    }
    p->grid_ = persistent_grid(form);
    *out = p;
    return SGEMM_STATUS_SUCCESS;
}

void GroupedPlan::release() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (uploaded_) (void)hipEventDestroy(uploaded_);
    if (dev_) (void)hipFree(dev_);
    if (staging_) (void)hipHostFree(staging_);
    exec_ = nullptr;
    graph_ = nullptr;
    uploaded_ = nullptr;
    dev_ = nullptr;
    staging_ = nullptr;
}

int GroupedPlan::set(int count, const sgemm_problem* problems, hipStream_t stream,
                     const int* forced_cfg) {
    if (count < 0 || count > capacity_ || (count > 0 && problems == nullptr))
        return SGEMM_STATUS_INVALID_VALUE;
    for (int i = 0; i < count; ++i) {
        int st = validate_problem(problems[i]);
        if (st != SGEMM_STATUS_SUCCESS) return st;
        if (forced_cfg && !grouped_has(form_, forced_cfg[i])) return SGEMM_STATUS_NOT_SUPPORTED;
    }
    std::vector<Problem> ps;
    std::vector<long long> cost;
    ps.reserve(count);
    flops_ = 0.0;
    for (int i = 0; i < count; ++i) {
        const sgemm_problem& q = problems[i];
        if (q.m == 0 || q.n == 0) continue;  // nothing to compute or store
        Problem p = to_problem(q);
        p.cfg = forced_cfg ? forced_cfg[i] : choose_grouped_config(form_, q.m, q.n, q.k);
        const ConfigInfo& ci = config_info(form_, p.cfg);
        p.tiles_m = div_up(q.m, ci.bm);
        p.tiles_n = div_up(q.n, ci.bn);
        p.reserved = i;  // submission index, for the stable order
        ps.push_back(p);
        cost.push_back((long long)std::max(p.k, 1) * ci.bm * ci.bn);
        flops_ += 2.0 * q.m * q.n * (double)p.k;
    }
    std::vector<int> order(ps.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return cost[a] > cost[b]; });
    // The previous upload must have left the staging buffer before we refill it.
    if (hipEventSynchronize(uploaded_) != hipSuccess) return SGEMM_STATUS_HIP_ERROR;
    GroupHeader* h = reinterpret_cast<GroupHeader*>(staging_);
    Problem* d = reinterpret_cast<Problem*>(static_cast<char*>(staging_) + sizeof(GroupHeader));
    long long tiles = 0;
    for (size_t r = 0; r < order.size(); ++r) {
        Problem p = ps[order[r]];
        p.tile_begin = (int)tiles;
        tiles += (long long)p.tiles_m * p.tiles_n;
        if (tiles > 0x7fffffffLL) return SGEMM_STATUS_INVALID_VALUE;
        d[r] = p;
    }
    h->count = (int)order.size();
    h->total_tiles = (int)tiles;
    count_ = h->count;
    total_tiles_ = h->total_tiles;
    const size_t bytes = sizeof(GroupHeader) + sizeof(Problem) * order.size();
    (void)hipGetLastError();
    if (hipMemcpyAsync(dev_, staging_, bytes, hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipEventRecord(uploaded_, stream) != hipSuccess)
        return SGEMM_STATUS_HIP_ERROR;
    return SGEMM_STATUS_SUCCESS;
}

int GroupedPlan::launch(hipStream_t stream) const {
    GroupHeader* hdr = reinterpret_cast<GroupHeader*>(dev_);
    const Problem* probs =
        reinterpret_cast<const Problem*>(static_cast<const char*>(dev_) + sizeof(GroupHeader));
    (void)hipGetLastError();
    if (form_ == kFormF32)
        hipLaunchKernelGGL(k_grouped<kFormF32>, dim3(grid_), dim3(256), 0, stream, hdr, probs);
    else
        hipLaunchKernelGGL(k_grouped<kFormBf16x3>, dim3(grid_), dim3(256), 0, stream, hdr, probs);
    return hipGetLastError() == hipSuccess ? SGEMM_STATUS_SUCCESS : SGEMM_STATUS_HIP_ERROR;
}

int GroupedPlan::run(hipStream_t stream) const {
    if (total_tiles_ == 0) return SGEMM_STATUS_SUCCESS;
    return launch(stream);
}

int GroupedPlan::capture(hipStream_t stream) {
    if (exec_) {
        (void)hipGraphExecDestroy(exec_);
        exec_ = nullptr;
    }
    if (graph_) {
        (void)hipGraphDestroy(graph_);
        graph_ = nullptr;
    }
    (void)hipGetLastError();
    if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return SGEMM_STATUS_HIP_ERROR;
    int st = launch(stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(stream, &g);
    if (st != SGEMM_STATUS_SUCCESS || e != hipSuccess || g == nullptr) {
        if (g) (void)hipGraphDestroy(g);
        return SGEMM_STATUS_HIP_ERROR;
    }
    graph_ = g;
    if (hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0) != hipSuccess) {
        exec_ = nullptr;
        return SGEMM_STATUS_HIP_ERROR;
    }
    return SGEMM_STATUS_SUCCESS;
}

int GroupedPlan::replay(hipStream_t stream) const {
    if (exec_ == nullptr) return SGEMM_STATUS_WRONG_STATE;
    return hipGraphLaunch(exec_, stream) == hipSuccess ? SGEMM_STATUS_SUCCESS
                                                        : SGEMM_STATUS_HIP_ERROR;
}

}  // namespace sgemm_amd

// ----------------------------------------------------------------- C ABI
using sgemm_amd::GroupedPlan;

extern "C" {

int sgemm_grouped_plan_create(sgemm_grouped_plan* plan, int capacity, int form) {
    if (plan == nullptr) return SGEMM_STATUS_INVALID_VALUE;
    GroupedPlan* p = nullptr;
    int st = GroupedPlan::create(&p, capacity, form);
    *plan = reinterpret_cast<sgemm_grouped_plan>(p);
    return st;
}

int sgemm_grouped_plan_set(sgemm_grouped_plan plan, int count, const sgemm_problem* problems,
                           sgemm_stream_t stream) {
    if (plan == nullptr) return SGEMM_STATUS_INVALID_VALUE;
    return reinterpret_cast<GroupedPlan*>(plan)->set(count, problems,
                                                     static_cast<hipStream_t>(stream), nullptr);
}

int sgemm_grouped_plan_set_configs(sgemm_grouped_plan plan, int count,
                                   const sgemm_problem* problems, const int* config_ids,
                                   sgemm_stream_t stream) {
    if (plan == nullptr || (count > 0 && config_ids == nullptr)) return SGEMM_STATUS_INVALID_VALUE;
    return reinterpret_cast<GroupedPlan*>(plan)->set(count, problems,
                                                     static_cast<hipStream_t>(stream), config_ids);
}

int sgemm_grouped_plan_run(sgemm_grouped_plan plan, sgemm_stream_t stream) {
    if (plan == nullptr) return SGEMM_STATUS_INVALID_VALUE;
    return reinterpret_cast<GroupedPlan*>(plan)->run(static_cast<hipStream_t>(stream));
}

int sgemm_grouped_plan_capture(sgemm_grouped_plan plan, sgemm_stream_t stream) {
    if (plan == nullptr) return SGEMM_STATUS_INVALID_VALUE;
    return reinterpret_cast<GroupedPlan*>(plan)->capture(static_cast<hipStream_t>(stream));
}

int sgemm_grouped_plan_replay(sgemm_grouped_plan plan, sgemm_stream_t stream) {
    if (plan == nullptr) return SGEMM_STATUS_INVALID_VALUE;
    return reinterpret_cast<GroupedPlan*>(plan)->replay(static_cast<hipStream_t>(stream));
}

int sgemm_grouped_plan_info(sgemm_grouped_plan plan, int* count, int* total_tiles, int* grid,
                            double* flops) {
    if (plan == nullptr) return SGEMM_STATUS_INVALID_VALUE;
    const GroupedPlan* p = reinterpret_cast<const GroupedPlan*>(plan);
    if (count) *count = p->count();
    if (total_tiles) *total_tiles = p->total_tiles();
    if (grid) *grid = p->grid();
    if (flops) *flops = p->flops();
    return SGEMM_STATUS_SUCCESS;
}

int sgemm_grouped_plan_destroy(sgemm_grouped_plan plan) {
    if (plan == nullptr) return SGEMM_STATUS_INVALID_VALUE;
    GroupedPlan* p = reinterpret_cast<GroupedPlan*>(plan);
    (void)hipDeviceSynchronize();
    p->release();
    delete p;
    return SGEMM_STATUS_SUCCESS;
}

int sgemm_grouped(int count, const sgemm_problem* problems, int form, sgemm_stream_t stream) {
    if (count < 0 || (count > 0 && problems == nullptr)) return SGEMM_STATUS_INVALID_VALUE;
    if (count == 0) return SGEMM_STATUS_SUCCESS;
    GroupedPlan* p = nullptr;
    int st = GroupedPlan::create(&p, count, form);
    if (st != SGEMM_STATUS_SUCCESS) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    st = p->set(count, problems, s, nullptr);
    if (st == SGEMM_STATUS_SUCCESS) st = p->run(s);
    if (hipStreamSynchronize(s) != hipSuccess && st == SGEMM_STATUS_SUCCESS)
        st = SGEMM_STATUS_HIP_ERROR;
    p->release();
    delete p;
    return st;
}

}  // extern "C"
