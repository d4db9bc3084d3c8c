// kernels.hpp -- internal interface between the host dispatcher (api.cpp) and
// the HIP translation units that hold the kernels.
#pragma once

#include "common.hpp"

namespace sgemm_amd {

enum Form { kFormF32 = 0, kFormBf16x3 = 1 };

struct ConfigInfo {
    int bm, bn, bk;  // workgroup tile
    int wm, wn;      // waves along m and n
    const char* name;
    int splitk;      // 1: the k range is split across workgroups (splitk_count);
                     // 2: only the tiles of a partial last round are split;
                     // 3: stream-K over the tiles of a partial last round
};

// Compute units of the current device (queried once per process).
int device_cu_count();

// Where a launch takes its split-K partials and arrival counters from when
// it is not the library's per-stream workspace: a fixed arena that never
// grows (the no-allocation entry point's static array).  Its counters are
// zero between launches, like the workspace's.
struct Arena {
    float* partials;
    size_t partial_bytes;
    int* counters;
    int counter_count;
};

// Number of k-splits for a tiles-sized grid, each split keeping at least
// min_ktiles k-tiles, at most 16.  BF16x3 (f32 = false): split only a grid of
// fewer than two workgroups per CU, up to two per CU.  Plain FP32: the split
// count nearest to four workgroups per CU -- its small tiles hide latency
// only with several workgroups per CU.
int splitk_count(long long tiles, int k, int bk, bool f32 = false, int min_ktiles = 4);

// Tail split: the r = tiles % slots tiles left after the full rounds of
// `slots` resident workgroups (per_cu per CU) are cut into s k-ranges.  The
// s chosen minimises the modelled time of what is left, in rounds of whole
// tiles: ceil(r s / slots) / s for the split units, plus the partials'
// traffic ((s + 1) partials of split_bytes through one workgroup at
// ~32 GB/s) over round_us, the time of one round.  Unsplit, the busiest CU
// holds ceil(r / CUs) of its per_cu workgroups, which run no faster each
// (the cost is that over per_cu).  At least four k-tiles per range, at most
// 16 ranges and 512 MiB of partials; 1 (and *cost = the unsplit cost) when
// no split beats the unsplit round by 3 %.
int tail_split_count(long long tiles, long long slots, int per_cu, int k, int bk,
                     size_t split_bytes, double round_us, double* cost);

// Useful BF16x3 rate of one CU running config 2's 256x256 tile (435 TFLOP/s
// over 256 CUs at 8192^3), for the length of a round in the tail-split model.
constexpr double kX3FlopsPerCuUs = 1.70e6;
// Partials of one 256x256 BF16x3 split: 128 FP32 accumulators x 512 threads.
constexpr size_t kX3TileSplitBytes = 256 * 256 * 4;
// Minimum k-tiles per split of BF16x3 config 10 (the 128x128 tile of config
// 8 with split-K): 1536^3 116 -> 179, 2048^3 218 -> 241 TFLOP/s, while four
// k-tiles per split ran 1024^3 at half the 64x64 split config's speed.
constexpr int kX3WideSplitKtiles = 16;
// Stream-K (BF16x3 config 11): the cost of one piece of a cut tile, in
// k-tile iterations of the 256x256 tile -- writing its partials and, for the
// last piece, reading every piece back.  Fitted to 3072^3 (three pieces per
// tile) and 3584^3 (fifteen).
constexpr double kX3StreamKPieceIters = 7.0;

// Configs that balance a partial last round of tiles over the whole GPU --
// the tail split (9) and stream-K (11) of either form -- and BF16x3's packed
// configs (12, 13, which hold a pack workspace of the operands' size), mapped to
// the config that runs the same tile whole from the operands themselves
// (FP32 5, BF16x3 2; for config 12 also the same result bits); -1 for any
// other config.
constexpr int whole_tile_twin(int form, int config) {
    if (form == kFormBf16x3)
        return (config == 9 || config == 11 || config == 12 || config == 13) ? 2 : -1;
    if (config == 16) return 15;
    return (config == 9 || config == 11) ? 5 : -1;
}

// The same for plain FP32 config 11 (config 5's 128x128x16 tile, three
// workgroups per CU; 64 KB of partials per piece).
constexpr double kF32StreamKPieceIters = 4.0;

// Stream-K over the `rem` tiles of a partial round (or all tiles, without a
// full round) of `iters` k-tiles each, spread over `slots` workgroups:
// returns the k-tiles per workgroup w that minimises w + kPieceIters x
// (pieces of the most-cut tile, (iters - 1) / w + 2), at least the even share
// and at most 16 pieces and 512 MiB of partials (pieces x split_bytes per
// workgroup).  *part: the stream-K part's time in rounds of whole tiles.
// 0 when no w qualifies.
long long streamk_iters(long long rem, long long slots, long long iters, size_t split_bytes,
                        double piece_iters, double* part);

// Plans stream-K over the tail of a grid of `tiles` tiles on `slots`
// resident workgroups: sets sk->iters, sk->pieces and the workspace (partials
// and counters for *nwg stream-K workgroups) and returns in *nwg the number of
// stream-K workgroups, 0 when the launch should run whole tiles only (no
// partial round, one stream-K workgroup would suffice, or the stream is being
// captured and its workspace would have to grow).
hipError_t streamk_plan(hipStream_t stream, long long tiles, long long slots, int k, int bk,
                        size_t split_bytes, double piece_iters, SplitK* sk, long long* nwg,
                        const Arena* arena = nullptr);
// Plain FP32 rate of one CU running config 5's tile three at a time (141
// TFLOP/s over 256 CUs at 8192^3).
constexpr double kF32FlopsPerCuUs = 0.55e6;

// Split plan for a tiles-sized grid of BK-deep tiles whose split partials
// take split_bytes per tile: sets sk->splits (1 = no split) and kchunk, and
// for a split grid the workspace pointers.
hipError_t splitk_plan(hipStream_t stream, long long tiles, int k, int bk, size_t split_bytes,
                       SplitK* sk, bool f32 = false, int min_ktiles = 4,
                       const Arena* arena = nullptr);

// Split-K partials + zeroed per-tile counters for a launch on `stream`.
// arena == nullptr: the stream's workspace, grown on demand (the first call on
// a stream, or a larger problem, allocates; a stream under capture returns
// hipErrorStreamCaptureUnsupported instead of growing).  Otherwise the
// arena's memory, or hipErrorOutOfMemory when the launch needs more than it
// holds; nothing is allocated.
hipError_t splitk_workspace(hipStream_t stream, size_t partial_bytes, int tiles,
                            float** partials, int** counters, const Arena* arena = nullptr);

// Frees the stream's workspace (after waiting for the stream).
hipError_t splitk_workspace_release(hipStream_t stream);

// The stream's buffer for BF16x3 config 12's packed operand planes, at least
// `bytes` long: grown on demand like the split workspace (outgrown buffers
// are kept for captured graphs until release), never under capture
// (hipErrorStreamCaptureUnsupported), hipErrorOutOfMemory when the
// allocation fails.  Its contents do not outlive a launch.
hipError_t pack_workspace(hipStream_t stream, size_t bytes, void** out);

// Configs the persistent grouped kernel carries (grouped.hip).  Its
// registers and LDS are the maximum over these, which sets its occupancy:
// FP32 {1, 2, 5} fit three workgroups per CU, BF16x3 {1, 8} two.
constexpr bool grouped_config(int form, int cfg) {
    return form == kFormF32 ? (cfg == 1 || cfg == 2 || cfg == 5) : (cfg == 1 || cfg == 8);
}

// Plain FP32 configs (sgemm_f32.hip).
int f32_config_count();
const ConfigInfo& f32_config(int id);
// arena: see splitk_workspace (nullptr: the stream's workspace).
hipError_t launch_f32(int id, const Problem& p, hipStream_t stream, const Arena* arena = nullptr);

// Plain FP32 config 15 (f32_256.hip): the 256x256x32 tile of four waves,
// one per SIMD, for 16-byte aligned A and B columns and k a multiple of 32
// (at least 64); config 5's result bits.  launch_f32 runs config 9 where
// f32_256_supported does not hold.
// Config 16 is config 15 with a stream-K tail (streamk): the tiles of a
// partial last round are spread k-tile by k-tile over the CUs, as config 11
// does for config 5's tile; with an arena it runs config 15's whole tiles.
bool f32_256_supported(const Problem& p);
hipError_t launch_f32_256(const Problem& p, hipStream_t stream, bool streamk,
                          const Arena* arena = nullptr);
// Config 15/16 cost model (8192x8192x{8192, 16384} on one box): one
// 256x256x32 k-tile of one tile takes kF256KtileUs on its CU (152 TFLOP/s
// over 256 CUs), a round of tiles kF256RoundUs more; one stream-K piece of
// config 16 costs kF256StreamKPieceIters k-tiles (storing 256 KB of
// partials and, for the last piece, reading every piece back; fitted to
// 9216^3, 10240^3 and 13312^3).
constexpr double kF256KtileUs = 7.06;
constexpr double kF256RoundUs = 25.0;
constexpr double kF256StreamKPieceIters = 4.0;
constexpr size_t kF256SplitBytes = 64 * 256 * 16;  // accumulators of one piece

// BF16x3 configs (sgemm_bf16x3.hip).
int bf16x3_config_count();
const ConfigInfo& bf16x3_config(int id);
hipError_t launch_bf16x3(int id, const Problem& p, hipStream_t stream,
                         const Arena* arena = nullptr);
// BF16x3 configs 12 and 13 (packed_x3.hip): split once into packed planes,
// then the packed 256x256 tile, whole (12) or with config 11's stream-K tail
// (13); where the planes cannot be had (capture, arena, size) config 2 or 11
// runs instead, with the same bits.
hipError_t launch_x3_packed(const Problem& p, hipStream_t stream, const Arena* arena,
                            bool streamk);
// Packed-plane rate of config 12's k loop relative to config 2's (GEMM part,
// 8192^3 on one box: 526-534 against 433-442 TFLOP/s) and the pack's read +
// write rate in bytes per microsecond (4.9-5.0 TB/s at 8192^2), for the cost
// model.
constexpr double kX3PackedRate = 1.22;
constexpr double kX3PackBytesPerUs = 5.0e6;

// Naive one-thread-per-output kernel (sgemm_f32.hip).
hipError_t launch_basic(const Problem& p, hipStream_t stream);
// The same with BLAS beta semantics, for leading dimensions above kMaxLd.
hipError_t launch_wide(const Problem& p, hipStream_t stream);

// C := beta*C for the k == 0 / alpha == 0 shortcut (sgemm_f32.hip).
hipError_t launch_scale_c(const Problem& p, hipStream_t stream);

// The column-major argument rules of the single-problem entry points
// (api.cpp): SGEMM_STATUS_SUCCESS or SGEMM_STATUS_INVALID_VALUE.
int validate_args(int m, int n, int k, const float* alpha, const void* A, int lda,
                  const void* B, int ldb, const float* beta, const void* C, int ldc);

// The shared body of the single-problem entry points (api.cpp): validation,
// the k == 0 / alpha == 0 shortcut, the wide-ld fallback and the config
// launch; config_id < 0 picks from the autotune table.  Returns an
// SGEMM_STATUS_* code; with an arena, SGEMM_STATUS_SCRATCH_TOO_SMALL when
// the chosen plan needs more than the arena holds (nothing is launched).
int run_problem(int form, int config_id, int m, int n, int k, const float* alpha,
                const float* A, int lda, const float* B, int ldb, const float* beta, float* C,
                int ldc, hipStream_t stream, const Arena* arena);

inline int config_count(int form) {
    return form == kFormF32 ? f32_config_count() : form == kFormBf16x3 ? bf16x3_config_count() : 0;
}
inline const ConfigInfo& config_info(int form, int id) {
    return form == kFormF32 ? f32_config(id) : bf16x3_config(id);
}

// Operands must be 4-byte aligned (every float pointer is): the tiles load
// and store four floats at a time with 16-byte accesses, which need dword
// alignment only, whatever the leading dimensions.
inline bool dword_aligned(const void* q) { return (reinterpret_cast<uintptr_t>(q) & 3u) == 0; }

// Every column of A and B starts on a 16-byte boundary: the condition for the
// FP32 direct-to-LDS tile (configs 5, 9 and 10 run it when this holds).
inline bool columns_16b_aligned(const Problem& p) {
    return ((reinterpret_cast<uintptr_t>(p.A) | reinterpret_cast<uintptr_t>(p.B)) & 15u) == 0 &&
           p.lda % 4 == 0 && p.ldb % 4 == 0;
}

}  // namespace sgemm_amd
