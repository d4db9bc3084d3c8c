// autotune.hip -- the project's autotuner: picks a tile config per
// (form, size bucket of m, n, k) by timing the candidates on the device, keeps
// the winners in a table, and saves / loads that table as JSON.
//
// Selection is a pure function of (form, grouped, m, n, k) and the loaded
// table: the same shape always gets the same config, which is what makes a
// problem's bits independent of what else runs beside it (grouped launches,
// the serving loop).  Without a table entry the built-in heuristic applies.
// A table named by the environment variable SGEMM_AMD_TUNE_TABLE is loaded on
// first use.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "sgemm.h"
#include "autotune.hpp"
#include "grouped.hpp"
#include "kernels.hpp"

namespace sgemm_amd {

namespace {

struct Entry {
    int config;
    float tflops;
};

struct Table {
    std::mutex mu;
    std::map<int, Entry> entries;  // key(form, grouped, mb, nb, kb)
    bool env_loaded = false;
};

Table& table() {
    static Table t;
    return t;
}

int key(int form, int grouped, int mb, int nb, int kb) {
    return (((form * 2 + grouped) * 16 + mb) * 16 + nb) * 16 + kb;
}

void unkey(int key, int& form, int& grouped, int& mb, int& nb, int& kb) {
    kb = key % 16;
    key /= 16;
    nb = key % 16;
    key /= 16;
    mb = key % 16;
    key /= 16;
    grouped = key % 2;
    form = key / 2;
}

bool grouped_capable(int form, int cfg) { return grouped_config(form, cfg); }

int load_locked(Table& t, const char* path);

void ensure_env_table(Table& t) {
    if (t.env_loaded) return;
    t.env_loaded = true;
    const char* p = std::getenv("SGEMM_AMD_TUNE_TABLE");
    if (p && *p) load_locked(t, p);
}

int heuristic(int form, int grouped, int m, int n, int k, bool aligned) {
    if (form == kFormF32) {
        if (grouped) {
            if (m >= 96 && n >= 96) return 5;
            return m >= 96 ? 2 : 1;
        }
        // Measured on MI355X (square sweeps 1024..8192 step 256, with
        // aligned and with odd leading dimensions, the sizes whose last
        // round is short, and the odd shapes).  From about one round of
        // 128x128 tiles on, config 5's tile is the fastest, with its tail
        // split (config 9) cutting k for the tiles of a partial last round
        // where tail_split_count's model says it pays: 4352^3 128 -> 140,
        // 4608^3 122 -> 139, 5888^3 133 -> 145, 7168^3 137 -> 146 TFLOP/s
        // (direct-loaded tile, 1024 resident).  The register-staged tile
        // (768 resident) takes config 9 from 700 tiles: 3584^3 105 -> 137.
        // Short k (<= 256) is bound by reading and writing C: 128x64 with
        // C prefetched at the start of each tile (config 14): 8192x8192x127
        // (alpha 1.5, beta 0.5) 76 -> 98, 8192x8192x256 109 -> 116 TFLOP/s
        // over config 2.  With aligned columns and a full round of tiles the
        // direct-loaded 128x128 tile (config 9) overtakes it above k = 160:
        // 8192x8192x192 111 -> 114, x256 116 -> 123, 16384x16384x192 111 ->
        // 122 (at k = 160 they tie, at 128 config 14 leads 101 to 96).
        const long long t128 = (long long)div_up(m, 128) * div_up(n, 128);
        const int short_k = aligned && t128 >= 1024 ? 160 : 256;
        if (k <= short_k && t128 >= 512) return 14;
        if (aligned) {
            // Configs 15 and 16 (256x256x32, one wave per SIMD and one
            // workgroup per CU; 15 gives config 5's bits, 16 adds a stream-K
            // tail): 8192^3 145.4 -> 150.1, 13312^3 146.3 -> 150.1 (16),
            // 10240^3 145.8 -> 149.1 (16), 8192x1024x8192 135 -> 141 (16)
            // TFLOP/s on one box.  Modelled time t256 (microseconds): a round
            // of their tiles costs I k-tiles of kF256KtileUs each plus
            // kF256RoundUs (prologue and epilogue of the whole chip at once),
            // config 16's partial round what streamk_iters charges.  It is
            // weighed against the time the configs below would take, with 3 %
            // in its favour: against measurements the model overstates it
            // (4096^2x2048 141.3 vs config 9's 138.5, 8192x1024x8192 141.0 vs
            // config 5's 135) more often than it understates it (8192^2x2048
            // 141.8 vs 143.0, 9216^3 144.5 vs 145.6).  Below k = 2048 the
            // round cost makes config 9 faster (8192^2x512 117 vs 130).
            double t256 = 0.0;
            int c256 = -1;
            if (k % 32 == 0 && k >= 2048) {
                const long long cus = device_cu_count();
                const long long tiles = (long long)div_up(m, 256) * div_up(n, 256);
                const long long iters = k / 32, full = tiles / cus, rem = tiles % cus;
                const double round = (double)iters * kF256KtileUs + kF256RoundUs;
                const double t15 = (double)(full + (rem > 0 ? 1 : 0)) * round;
                double t16 = t15;
                double part = 1.0;
                if (rem > 0 && streamk_iters(rem, cus, iters, kF256SplitBytes,
                                             kF256StreamKPieceIters, &part) > 0)
                    t16 = (double)full * round + part * iters * kF256KtileUs + kF256RoundUs;
                c256 = t16 < t15 ? 16 : 15;
                t256 = std::min(t15, t16) / 1.03;
            }
            const double flops_us = 2.0 * m * n * (double)k / 1e6;  // MFLOP
            if (t128 >= 1024) {
                // config 9: its padded 128x128 tiles at 145 TFLOP/s (its tail
                // split fills the last round)
                const double t9 = 2.0 * t128 * 16384.0 * k / 145.0e6;
                return (c256 >= 0 && t256 < t9) ? c256 : 9;
            }
            // Short of one round of 1024 direct-loaded 128x128 tiles, with k
            // >= 1024 and from 180 tiles: every tile of a config takes the
            // same time, so with c tiles per CU on average the busiest CU
            // runs ceil(c) and a config reaches its full-round rate times
            // c / ceil(c).  Full-round rates measured on 2048..4096 step 128:
            // 128x128 (config 5) 141 TFLOP/s (0.83 of it with one tile per
            // CU), 64x64 (config 7) 133, config 5's tile with split-K
            // (config 10, splits to fill four workgroups per CU) 131 per
            // split unit.  The largest modelled rate wins: within 2.5 % of
            // the best config at every size of that sweep, where a fixed
            // rule lost 7-10 % at 2432, 2944, 3456 and 3968 (e.g. 3968^3
            // 123 -> 132, 2944^3 114 -> 123 TFLOP/s).
            if (k >= 1024 && t128 >= 180) {
                const double cus = device_cu_count();
                auto fill = [](double c) { return c / std::ceil(c); };
                const double c5 = (double)t128 / cus;
                const double e5 = 141.0 * fill(c5) * (c5 <= 1.0 ? 0.83 : 1.0);
                const double c7 = (double)div_up(m, 64) * div_up(n, 64) / cus;
                const double e7 = 133.0 * fill(c7);
                const long long s10 = std::llround(4.0 * cus / (double)t128);
                const double e10 = s10 >= 2 ? 131.0 * fill((double)(t128 * s10) / cus) : 0.0;
                const double e = std::max({e5, e7, e10});
                // (only for long k here: at k = m the piece costs weigh more
                // than the model gives them, 3072^3 124 vs config 7's 133)
                if (c256 >= 0 && k >= 4096 && t256 < flops_us / e) return c256;
                if (e5 >= e7 && e5 >= e10) return 5;
                return e10 > e7 ? 10 : 7;
            }
            // 700-1023 tiles with k < 1024: 64x64 tiles (3584x3584x512 104
            // -> 114, 4096x3584x768 121 -> 125).
            if (t128 >= 700) return 7;
        } else if (t128 >= 700) {
            return 9;
        }
        // Below that with k >= 1024, from 180 tiles of 128x128: with t 64x64
        // tiles per CU on average the busiest CU runs ceil(t).  Where that
        // loads the CUs unevenly, config 5's tile with split-K (config 10)
        // is faster, else 64x64 tiles (config 7: loads two k-tiles ahead,
        // split-K when the grid holds fewer than about four workgroups per
        // CU).  The threshold is 0.9 below 350 tiles of 128x128, 0.95 from
        // there: 1792^3 (3.1 64x64 tiles per CU) 94 -> 102, 2304^3 (5.1)
        // 106 -> 124, 2560^3 (6.25) 114 -> 124, 2816^3 (7.6) 122 -> 131
        // TFLOP/s; while 3072^3 (9.0) 128 vs 124 and 3328^3 (10.6) 125 vs
        // 124 keep 64x64 (stream-K, config 11, ran 116-123 on these).
        if (k >= 1024 && t128 >= 180) {
            const double t = (double)div_up(m, 64) * div_up(n, 64) / device_cu_count();
            if (t / std::ceil(t) < (t128 >= 350 ? 0.95 : 0.9)) return 10;
            if (t128 >= 350) return 7;
        }
        // One 64x64 tile per CU (90-100 % of the CUs) and k <= 1536: the
        // k-group tile (config 12) over split-K: 1024^3 86 -> 92, 1000^3 79
        // -> 84 TFLOP/s (at k = 4096 split-K wins, 114 vs 105).
        const long long t64 = (long long)div_up(m, 64) * div_up(n, 64);
        const long long cus = device_cu_count();
        if (k <= 1536 && t64 <= cus && 10 * t64 >= 9 * cus) return 12;
        return 7;
    }
    if (grouped) return (m >= 96 && n >= 96) ? 8 : 1;
    // BF16x3: configs 2 and 3 hold one workgroup per CU, 5, 7 (64x64 with
    // split-K) and 8 (128x128) two.  Every tile of a config takes the same
    // time, so a launch runs in rounds of CUs x per_cu workgroups.  A full
    // round costs per_cu x (tile area / splits) / rate, rate being the
    // config's relative speed measured at 8192^3.  In a last, partial round
    // a CU holds w < per_cu workgroups, which then run faster each: a lone
    // one of a two-per-CU config gets `lone` of the CU -- about 60 % for
    // configs 5 and 7, 78 % for 8 and 10, whose k loop overlaps the split
    // with the MFMAs inside each wave (2048^3 on config 8: 253 TFLOP/s with
    // one workgroup per CU against 326 at 8192^3).  The cheapest config wins.  Configs 7 and 10
    // split k when their grid is short of two workgroups per CU
    // (splitk_count); 10 only into splits of at least 16 k-tiles.
    // Config 9 is config 2 with a tail split: from one full round on, the
    // tiles of a partial last round are cut into k-ranges, which costs what
    // tail_split_count models instead of a whole round; when it would not
    // split it is config 2 and is left to config 2.  Config 11 is config 2
    // with stream-K over those tiles (or over all tiles, short of a full
    // round), costing what streamk_iters models: 3072^3 244 -> 314, 5120^3
    // 352 -> 387 TFLOP/s.
    // Configs 12 and 13 are configs 2 and 11 on operands split once into
    // packed planes: their k loop runs kX3PackedRate times config 2's, plus
    // the pack pass (A and B read, their padded planes written, at
    // kX3PackBytesPerUs, and two launches more).
    enum { kWhole = 0, kSplitK = 1, kTailSplit = 2, kStreamK = 3 };
    struct Cand { int id, bm, bn, bk, per_cu; double rate; int split; double lone; bool packed; };
    static const Cand cands[] = {
        {2, 256, 256, 32, 1, 1.00, kWhole, 1.0},   {9, 256, 256, 32, 1, 1.00, kTailSplit, 1.0},
        {5, 256, 128, 16, 2, 0.81, kWhole, 0.6},   {8, 128, 128, 32, 2, 0.78, kWhole, 0.78},
        {3, 256, 128, 32, 1, 0.62, kWhole, 1.0},   {7, 64, 64, 64, 2, 0.43, kSplitK, 0.6},
        {10, 128, 128, 32, 2, 0.66, kSplitK, 0.78},  // config 8's rate less its split's partials
        {11, 256, 256, 32, 1, 1.00, kStreamK, 1.0},
        {12, 256, 256, 32, 1, kX3PackedRate, kWhole, 1.0, true},
        {13, 256, 256, 32, 1, kX3PackedRate, kStreamK, 1.0, true},
    };
    const long long cus = device_cu_count();
    int best = 7;
    double best_cost = 0;
    for (const Cand& c : cands) {
        const long long tiles = (long long)div_up(m, c.bm) * div_up(n, c.bn);
        const long long slots = cus * c.per_cu;
        double cost;
        if (c.split == kTailSplit) {
            double tail = 1.0;
            const int s = tail_split_count(tiles, slots, c.per_cu, k, c.bk, kX3TileSplitBytes,
                                           2.0 * c.bm * c.bn * k / kX3FlopsPerCuUs, &tail);
            if (s <= 1) continue;
            const double round = (double)c.per_cu * c.bm * c.bn / c.rate;
            cost = (double)(tiles / slots) * round + tail * round;
        } else if (c.split == kStreamK) {
            double part = 1.0;
            if (tiles % slots == 0 ||
                streamk_iters(tiles % slots, slots, div_up(k, c.bk), kX3TileSplitBytes,
                              kX3StreamKPieceIters, &part) <= 0)
                continue;
            // The stream-K part replaces a round of whole tiles; take it only
            // where the model saves at least 5 % of that round (3584^3, 196
            // tiles: modelled 4.5 % faster than whole tiles, measured 2 %
            // slower).
            if (part > 0.95) continue;
            const double round = (double)c.per_cu * c.bm * c.bn / c.rate;
            cost = (double)(tiles / slots) * round + part * round;
        } else {
            const long long splits =
                c.split == kSplitK
                    ? splitk_count(tiles, k, c.bk, false, c.id == 10 ? kX3WideSplitKtiles : 4)
                    : 1;
            const long long wgs = tiles * splits;
            // Config 10's partials cost about four k-tiles per split (written,
            // then read back by the last split): 2048^3 (32 k-tiles per
            // split) runs faster unsplit on config 8 (255 vs 245 TFLOP/s),
            // 4096x1024x4096 (64 per split) faster split (299 vs 282).
            const double split_cost =
                c.id == 10 && splits > 1 ? 1.0 + 4.0 / ((double)k / splits / c.bk) : 1.0;
            const double work = (double)c.bm * c.bn / splits / c.rate * split_cost;  // one workgroup, full CU share
            cost = (double)(wgs / slots) * c.per_cu * work;
            const long long rem = wgs % slots;
            if (rem > 0) {
                const long long w = div_up(rem, cus);
                const double u =
                    w >= c.per_cu ? 1.0 : c.lone + (1.0 - c.lone) * (double)(w - 1) / (c.per_cu - 1);
                cost += (double)w * work / u;
            }
        }
        if (c.packed) {
            // In the cost's unit (tile area at config 2's rate, 2 k /
            // kX3FlopsPerCuUs microseconds).
            const double kp = 32.0 * div_up(k, 32);
            const double bytes = 4.0 * ((double)m + n) * k +
                                 4.0 * (256.0 * div_up(m, 256) + 256.0 * div_up(n, 256)) * kp;
            cost += (bytes / kX3PackBytesPerUs + 5.0) / (2.0 * k / kX3FlopsPerCuUs);
        }
        if (best_cost == 0 || cost < best_cost) {
            best = c.id;
            best_cost = cost;
        }
    }
    return best;
}

int lookup(int form, int grouped, int m, int n, int k, bool aligned = true) {
    Table& t = table();
    std::lock_guard<std::mutex> lock(t.mu);
    ensure_env_table(t);
    auto it = t.entries.find(key(form, grouped, size_bucket(m), size_bucket(n), size_bucket(k)));
    if (it != t.entries.end()) {
        const int c = it->second.config;
        if (c >= 0 && c < config_count(form) && (!grouped || grouped_capable(form, c))) return c;
    }
    return heuristic(form, grouped, m, n, k, aligned);
}

int load_locked(Table& t, const char* path) {
    FILE* f = std::fopen(path, "r");
    if (!f) return SGEMM_STATUS_IO_ERROR;
    char line[512];
    std::map<int, Entry> fresh;
    bool ok = true;
    while (std::fgets(line, sizeof(line), f)) {
        if (!std::strstr(line, "\"form\"")) continue;
        int form, grouped, mb, nb, kb, cfg;
        float tf = 0.f;
        if (std::sscanf(line,
                        " {\"form\": %d, \"grouped\": %d, \"mb\": %d, \"nb\": %d, \"kb\": %d, "
                        "\"config\": %d, \"tflops\": %f",
                        &form, &grouped, &mb, &nb, &kb, &cfg, &tf) != 7 ||
            form < 0 || form > 1 || grouped < 0 || grouped > 1 || mb < 0 || mb > 15 || nb < 0 ||
            nb > 15 || kb < 0 || kb > 15) {
            ok = false;
            break;
        }
        fresh[key(form, grouped, mb, nb, kb)] = Entry{cfg, tf};
    }
    std::fclose(f);
    if (!ok) return SGEMM_STATUS_IO_ERROR;
    t.entries = fresh;
    return SGEMM_STATUS_SUCCESS;
}

// Uniform [-1, 1) from a hash of the index (tuning inputs only).
__global__ void k_fill_tune(float* x, size_t n, unsigned seed) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) {
        unsigned h = (unsigned)i * 2654435761u ^ seed;
        h ^= h >> 15;
        h *= 2246822519u;
        h ^= h >> 13;
        x[i] = (float)(h >> 8) * (2.0f / 16777216.0f) - 1.0f;
    }
}

}  // namespace

int size_bucket(int x) {
    if (x <= 1) return 0;
    int b = 0;
    unsigned v = (unsigned)(x - 1);
    while (v) {
        ++b;
        v >>= 1;
    }
    return b > 15 ? 15 : b;
}

int choose_config(int form, int m, int n, int k, bool aligned) {
    return lookup(form, 0, m, n, k, aligned);
}
int choose_grouped_config(int form, int m, int n, int k) { return lookup(form, 1, m, n, k); }

}  // namespace sgemm_amd

using namespace sgemm_amd;

extern "C" {

int sgemm_autotune_lookup(int form, int grouped, int m, int n, int k, int* config_id,
                          int* from_table) {
    if (form != kFormF32 && form != kFormBf16x3) return SGEMM_STATUS_NOT_SUPPORTED;
    if (m < 1 || n < 1 || k < 0 || config_id == nullptr) return SGEMM_STATUS_INVALID_VALUE;
    grouped = grouped ? 1 : 0;
    *config_id = lookup(form, grouped, m, n, k);
    if (from_table) {
        Table& t = table();
        std::lock_guard<std::mutex> lock(t.mu);
        *from_table = t.entries.count(key(form, grouped, size_bucket(m), size_bucket(n),
                                          size_bucket(k))) ? 1 : 0;
    }
    return SGEMM_STATUS_SUCCESS;
}

int sgemm_autotune_set(int form, int grouped, int m, int n, int k, int config_id) {
    if (form != kFormF32 && form != kFormBf16x3) return SGEMM_STATUS_NOT_SUPPORTED;
    if (m < 1 || n < 1 || k < 0) return SGEMM_STATUS_INVALID_VALUE;
    if (config_id < 0 || config_id >= config_count(form) ||
        (grouped && !grouped_capable(form, config_id)))
        return SGEMM_STATUS_NOT_SUPPORTED;
    Table& t = table();
    std::lock_guard<std::mutex> lock(t.mu);
    ensure_env_table(t);
    t.entries[key(form, grouped ? 1 : 0, size_bucket(m), size_bucket(n), size_bucket(k))] =
        Entry{config_id, 0.f};
    return SGEMM_STATUS_SUCCESS;
}

int sgemm_autotune_clear(void) {
    Table& t = table();
    std::lock_guard<std::mutex> lock(t.mu);
    t.entries.clear();
    t.env_loaded = true;  // an explicit clear also drops the environment table
    return SGEMM_STATUS_SUCCESS;
}

int sgemm_autotune_size(void) {
    Table& t = table();
    std::lock_guard<std::mutex> lock(t.mu);
    ensure_env_table(t);
    return (int)t.entries.size();
}

int sgemm_autotune_save(const char* path) {
    if (path == nullptr) return SGEMM_STATUS_INVALID_VALUE;
    Table& t = table();
    std::lock_guard<std::mutex> lock(t.mu);
    ensure_env_table(t);
    FILE* f = std::fopen(path, "w");
    if (!f) return SGEMM_STATUS_IO_ERROR;
    std::fprintf(f, "{\"format\": \"sgemm_amd autotune table v1\",\n \"entries\": [\n");
    size_t i = 0;
    for (const auto& kv : t.entries) {
        int form, grouped, mb, nb, kb;
        unkey(kv.first, form, grouped, mb, nb, kb);
        std::fprintf(f,
                     "  {\"form\": %d, \"grouped\": %d, \"mb\": %d, \"nb\": %d, \"kb\": %d, "
                     "\"config\": %d, \"tflops\": %.3f}%s\n",
                     form, grouped, mb, nb, kb, kv.second.config, kv.second.tflops,
                     ++i < t.entries.size() ? "," : "");
    }
    std::fprintf(f, " ]\n}\n");
    return std::fclose(f) == 0 ? SGEMM_STATUS_SUCCESS : SGEMM_STATUS_IO_ERROR;
}

int sgemm_autotune_load(const char* path) {
    if (path == nullptr) return SGEMM_STATUS_INVALID_VALUE;
    Table& t = table();
    std::lock_guard<std::mutex> lock(t.mu);
    t.env_loaded = true;
    return load_locked(t, path);
}

int sgemm_autotune_bucket(int form, int grouped, int m, int n, int k, int iters,
                          sgemm_stream_t stream, int* best_config, float* best_tflops) {
    if (form != kFormF32 && form != kFormBf16x3) return SGEMM_STATUS_NOT_SUPPORTED;
    if (m < 1 || n < 1 || k < 1 || iters < 1) return SGEMM_STATUS_INVALID_VALUE;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t na = (size_t)m * k, nb = (size_t)k * n, nc = (size_t)m * n;
    float *A = nullptr, *B = nullptr, *C = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int status = SGEMM_STATUS_SUCCESS;
    if (hipMalloc(&A, na * 4) != hipSuccess || hipMalloc(&B, nb * 4) != hipSuccess ||
        hipMalloc(&C, nc * 4) != hipSuccess) {
        status = SGEMM_STATUS_ALLOC_FAILED;
    } else if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        status = SGEMM_STATUS_HIP_ERROR;
    }
    int best = -1;
    float best_tf = 0.f;
    if (status == SGEMM_STATUS_SUCCESS) {
        hipLaunchKernelGGL(k_fill_tune, dim3(1024), dim3(256), 0, s, A, na, 1u);
        hipLaunchKernelGGL(k_fill_tune, dim3(1024), dim3(256), 0, s, B, nb, 2u);
        const float one = 1.f, zero = 0.f;
        const double flops = 2.0 * m * n * (double)k;
        // Grouped buckets are timed inside a grouped launch of `copies` copies
        // of the problem (each with its own C), enough to fill the device:
        // a grouped launch runs many problems at once, so what matters there
        // is a config's throughput per tile, not one problem's latency.
        int copies = 1;
        GroupedPlan* plan = nullptr;
        std::vector<sgemm_problem> probs;
        float* Cc = nullptr;
        if (grouped) {
            const long long t128 = (long long)div_up(m, 128) * div_up(n, 128);
            copies = (int)std::min<long long>(256, div_up(4LL * device_cu_count(), t128));
            while (copies > 1 && (size_t)copies * nc * 4 > ((size_t)512 << 20)) copies /= 2;
            if (hipMalloc(&Cc, (size_t)copies * nc * 4) != hipSuccess ||
                GroupedPlan::create(&plan, copies, form) != SGEMM_STATUS_SUCCESS) {
                status = SGEMM_STATUS_ALLOC_FAILED;
                copies = 0;
            }
            for (int i = 0; i < copies; ++i)
                probs.push_back(sgemm_problem{m, n, k, 1.f, 0.f, A, m, B, k, Cc + (size_t)i * nc, m});
        }
        for (int c = 0; copies > 0 && c < config_count(form); ++c) {
            if (grouped && !grouped_capable(form, c)) continue;
            auto once = [&]() -> int {
                if (!grouped) return sgemm_run_config(form, c, m, n, k, &one, A, m, B, k, &zero, C, m, s);
                return plan->run(s);
            };
            if (grouped) {
                std::vector<int> cfgs(copies, c);
                if (plan->set(copies, probs.data(), s, cfgs.data()) != SGEMM_STATUS_SUCCESS) continue;
            }
            // warm-up, then the timed repetitions
            int st = once();
            if (st != SGEMM_STATUS_SUCCESS) continue;
            if (hipEventRecord(e0, s) != hipSuccess) {
                status = SGEMM_STATUS_HIP_ERROR;
                break;
            }
            for (int r = 0; r < iters && st == SGEMM_STATUS_SUCCESS; ++r) st = once();
            float ms = 0.f;
            if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                status = SGEMM_STATUS_HIP_ERROR;
                break;
            }
            if (st != SGEMM_STATUS_SUCCESS || !(ms > 0.f)) continue;
            const float tf = (float)(flops * copies * iters / (ms * 1e-3) / 1e12);
            if (best < 0 || tf > best_tf) {
                best = c;
                best_tf = tf;
            }
        }
        if (plan) {
            (void)hipStreamSynchronize(s);
            plan->release();
            delete plan;
        }
        if (Cc) (void)hipFree(Cc);
    }
    // Teardown: failures here cannot change the result already measured.
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(A);
    (void)hipFree(B);
    (void)hipFree(C);
    if (status == SGEMM_STATUS_SUCCESS && best < 0) status = SGEMM_STATUS_HIP_ERROR;
    if (status == SGEMM_STATUS_SUCCESS) {
        Table& t = table();
        std::lock_guard<std::mutex> lock(t.mu);
        ensure_env_table(t);
        t.entries[key(form, grouped ? 1 : 0, size_bucket(m), size_bucket(n), size_bucket(k))] =
            Entry{best, best_tf};
        if (best_config) *best_config = best;
        if (best_tflops) *best_tflops = best_tf;
    }
    return status;
}

}  // extern "C"
