// sgemm_bf16x3.hip -- single-problem launches of the BF16x3 tile (tile_x3.hpp).
#include "tile_f32.hpp"
#include "tile_x3.hpp"

namespace sgemm_amd {

constexpr int kX3GroupN = 8;

// WPE: waves per SIMD the register allocation must allow (T::WPE).
// SPLIT: workgroups [0, sk.dp_tiles) compute tiles [0, sk.dp_tiles) whole;
// the rest are (tile, split) units of the remaining tiles, split index
// outermost (as sgemm_f32.hip).
template <class T, bool SPLIT>
__global__ __launch_bounds__(T::NT) __attribute__((amdgpu_waves_per_eu(T::WPE)))
void k_sgemm_bf16x3(Problem P, SplitK sk) {
    __shared__ __attribute__((aligned(16))) char smem[T::LDS_BYTES];
    int tm, tn;
    const int b = blockIdx.x;
    if (!SPLIT || b < sk.dp_tiles) {
        tile_coords(xcd_remap(b, SPLIT ? sk.dp_tiles : (int)gridDim.x), P.tiles_m, P.tiles_n,
                    kX3GroupN, tm, tn);
        X3TileRunner<T>::run(P, tm, tn, smem);
    } else {
        const int units = (int)gridDim.x - sk.dp_tiles;
        const int u = xcd_remap(b - sk.dp_tiles, units);
        const int count = units / sk.splits;  // split tiles in this launch
        const int split = u / count;
        const int slot = u - split * count;
        tile_coords(sk.dp_tiles + slot, P.tiles_m, P.tiles_n, kX3GroupN, tm, tn);
        X3TileRunner<T>::run(P, tm, tn, smem, &sk, split, slot);
    }
}

// Stream-K over the tail: workgroups [0, sk.dp_tiles) compute tiles [0,
// sk.dp_tiles) whole (every full round of resident workgroups); workgroup
// dp_tiles + g of the rest runs k-tile iterations [g * iters, (g + 1) *
// iters) of the remaining tiles, taken tile-major with k-tiles in order.  A
// tile wholly inside one workgroup's range is stored straight to C; a tile
// cut at a workgroup boundary is reduced from its pieces in k order by the
// last of them to finish.
template <class T>
__global__ __launch_bounds__(T::NT) __attribute__((amdgpu_waves_per_eu(T::WPE)))
void k_sgemm_bf16x3_sk(Problem P, SplitK sk) {
    __shared__ __attribute__((aligned(16))) char smem[T::LDS_BYTES];
    int tm, tn;
    const int b = blockIdx.x;
    if (b < sk.dp_tiles) {
        tile_coords(xcd_remap(b, sk.dp_tiles), P.tiles_m, P.tiles_n, kX3GroupN, tm, tn);
        X3TileRunner<T>::run(P, tm, tn, smem);
        return;
    }
    streamk_pieces(P, sk, b, T::BK, [&](int tile, const KPiece& pc) {
        tile_coords(tile, P.tiles_m, P.tiles_n, kX3GroupN, tm, tn);
        X3TileRunner<T>::run_piece(P, tm, tn, smem, &sk, pc);
    });
}

// Ids 6-7 run the tiles of ids 0-1 with split-K (splitk = 1); id 8 is a
// 128x128 tile with one accumulator set, two workgroups per CU; id 9 runs
// the tile of id 2 with a tail split (splitk = 2, as FP32 config 9); id 10
// runs the tile of id 8 with split-K, each split at least 16 k-tiles deep
// (shorter splits of this tile lose to the 64x64 one).  Id 11 runs the tile
// of id 2 with stream-K over the tiles of a partial last round (splitk = 3).
// Id 12 splits the operands once into packed planes and runs a 256x256 tile
// on them (packed_x3.hip), with id 2's result bits; id 13 adds id 11's
// stream-K tail to it, with id 11's bits.
static const ConfigInfo kBf16x3Configs[] = {
    {128, 128, 32, 2, 2, "bf16x3_128x128x32_w2x2_m16_dual"},
    {64, 64, 64, 2, 2, "bf16x3_64x64x64_w2x2_m16_dual"},
    {256, 256, 32, 2, 4, "bf16x3_256x256x32_w2x4_m16_single"},
    {256, 128, 32, 2, 2, "bf16x3_256x128x32_w2x2_m32_single"},
    {128, 128, 32, 2, 2, "bf16x3_128x128x32_w2x2_m32_dual"},
    {256, 128, 16, 2, 2, "bf16x3_256x128x16_w2x2_m32_single_mid"},
    {128, 128, 32, 2, 2, "bf16x3_128x128x32_w2x2_m16_dual_splitk", 1},
    {64, 64, 64, 2, 2, "bf16x3_64x64x64_w2x2_m16_dual_splitk", 1},
    {128, 128, 32, 2, 2, "bf16x3_128x128x32_w2x2_m16_single_2cu"},
    {256, 256, 32, 2, 4, "bf16x3_256x256x32_w2x4_m16_single_tailsplit", 2},
    {128, 128, 32, 2, 2, "bf16x3_128x128x32_w2x2_m16_single_2cu_splitk", 1},
    {256, 256, 32, 2, 4, "bf16x3_256x256x32_w2x4_m16_single_streamk", 3},
    {256, 256, 32, 2, 4, "bf16x3_256x256x32_w2x4_m16_single_packed"},
    {256, 256, 32, 2, 4, "bf16x3_256x256x32_w2x4_m16_single_packed_streamk", 3},
};

int bf16x3_config_count() { return (int)(sizeof(kBf16x3Configs) / sizeof(kBf16x3Configs[0])); }
const ConfigInfo& bf16x3_config(int id) { return kBf16x3Configs[id]; }

template <class T, bool ALLOW_SPLIT, int MIN_KTILES = 4>
static hipError_t launch_x3(const Problem& p, hipStream_t s, const Arena* arena) {
    Problem q = p;
    q.tiles_m = div_up(p.m, T::BM);
    q.tiles_n = div_up(p.n, T::BN);
    q.tile_begin = 0;
    const long long tiles = (long long)q.tiles_m * q.tiles_n;
    if (tiles <= 0) return hipSuccess;
    if (tiles > 0x7fffffffLL) return hipErrorInvalidValue;
    SplitK sk{nullptr, nullptr, 1, p.k, 0};
    if (ALLOW_SPLIT) {
        constexpr size_t acc_bytes = T::MF == 16 ? 16 : 64;
        hipError_t e = splitk_plan(s, tiles, p.k, T::BK, T::MB * T::NB * acc_bytes * T::NT, &sk,
                                   false, MIN_KTILES, arena);
        if (e != hipSuccess) return e;
    }
    const bool split = ALLOW_SPLIT && sk.splits > 1;
    constexpr bool S = ALLOW_SPLIT;
    const dim3 grid((unsigned)(tiles * sk.splits));
    (void)hipGetLastError();
    if (split) hipLaunchKernelGGL((k_sgemm_bf16x3<T, S>), grid, dim3(T::NT), 0, s, q, sk);
    else hipLaunchKernelGGL((k_sgemm_bf16x3<T, false>), grid, dim3(T::NT), 0, s, q, sk);
    return hipGetLastError();
}

// Workgroups of the tail-split kernel resident on the device at once.
template <class T>
static long long x3_resident_workgroups() {
    static const long long n = [] {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sgemm_bf16x3<T, true>,
                                                         T::NT, 0) != hipSuccess ||
            per_cu < 1) {
            (void)hipGetLastError();
            per_cu = 1;
        }
        return (long long)per_cu * device_cu_count();
    }();
    return n;
}

// Tail split: whole tiles for every full round of resident workgroups, and
// the r tiles of a partial last round cut into k-ranges (tail_split_count).
// With one workgroup per CU the tiles of a partial round run no faster than
// in a full one, so without the split a round with a few tiles costs as much
// as a full one.
template <class T>
static hipError_t launch_x3_tail(const Problem& p, hipStream_t s, const Arena* arena) {
    Problem q = p;
    q.tiles_m = div_up(p.m, T::BM);
    q.tiles_n = div_up(p.n, T::BN);
    q.tile_begin = 0;
    const long long tiles = (long long)q.tiles_m * q.tiles_n;
    if (tiles <= 0) return hipSuccess;
    if (tiles > 0x7fffffffLL) return hipErrorInvalidValue;
    const long long slots = x3_resident_workgroups<T>();
    const long long rem = tiles % slots;
    constexpr size_t acc_bytes = T::MF == 16 ? 16 : 64;
    const size_t split_bytes = (size_t)T::MB * T::NB * acc_bytes * T::NT;
    double tail_cost = 1.0;
    const int per_cu = (int)(slots / device_cu_count());
    const int splits =
        tail_split_count(tiles, slots, per_cu, p.k, T::BK, split_bytes,
                         per_cu * 2.0 * T::BM * T::BN * p.k / kX3FlopsPerCuUs, &tail_cost);
    SplitK sk{nullptr, nullptr, 1, p.k, 0};
    if (splits > 1) {
        sk.kchunk = div_up(div_up(p.k, splits), T::BK) * T::BK;
        sk.splits = div_up(p.k, sk.kchunk);
        if (sk.splits > 1) {
            const hipError_t e = splitk_workspace(s, (size_t)rem * sk.splits * split_bytes,
                                                  (int)rem, &sk.partials, &sk.counters, arena);
            if (e == hipErrorStreamCaptureUnsupported) {
                sk.splits = 1;
                sk.kchunk = p.k;
            } else if (e != hipSuccess) {
                return e;
            }
        }
    }
    (void)hipGetLastError();
    if (sk.splits > 1) {
        sk.dp_tiles = (int)(tiles - rem);
        const dim3 grid((unsigned)(tiles - rem + rem * sk.splits));
        hipLaunchKernelGGL((k_sgemm_bf16x3<T, true>), grid, dim3(T::NT), 0, s, q, sk);
    } else {
        const dim3 grid((unsigned)tiles);
        hipLaunchKernelGGL((k_sgemm_bf16x3<T, false>), grid, dim3(T::NT), 0, s, q, sk);
    }
    return hipGetLastError();
}

// Stream-K over the tail: the r = tiles % slots tiles after the full rounds
// (all tiles when there is no full round) are spread, k-tile by k-tile, over
// up to `slots` workgroups of w k-tiles each (streamk_iters: w trades the
// longer ranges against fewer pieces per cut tile).  Falls back to one launch of whole tiles when
// there is no partial round, or when the stream is being captured and its
// workspace would have to grow.
template <class T>
static hipError_t launch_x3_streamk(const Problem& p, hipStream_t s, const Arena* arena) {
    Problem q = p;
    q.tiles_m = div_up(p.m, T::BM);
    q.tiles_n = div_up(p.n, T::BN);
    q.tile_begin = 0;
    const long long tiles = (long long)q.tiles_m * q.tiles_n;
    if (tiles <= 0) return hipSuccess;
    if (tiles > 0x7fffffffLL) return hipErrorInvalidValue;
    const long long slots = x3_resident_workgroups<T>();
    const long long rem = tiles % slots;
    constexpr size_t acc_bytes = T::MF == 16 ? 16 : 64;
    const size_t split_bytes = (size_t)T::MB * T::NB * acc_bytes * T::NT;
    SplitK sk{nullptr, nullptr, 1, p.k, 0};
    long long nwg = 0;
    const hipError_t e =
        streamk_plan(s, tiles, slots, p.k, T::BK, split_bytes, kX3StreamKPieceIters, &sk, &nwg,
                     arena);
    if (e != hipSuccess) return e;
    (void)hipGetLastError();
    if (nwg > 0) {
        sk.dp_tiles = (int)(tiles - rem);
        const dim3 grid((unsigned)(tiles - rem + nwg));
        hipLaunchKernelGGL((k_sgemm_bf16x3_sk<T>), grid, dim3(T::NT), 0, s, q, sk);
    } else {
        const dim3 grid((unsigned)tiles);
        hipLaunchKernelGGL((k_sgemm_bf16x3<T, false>), grid, dim3(T::NT), 0, s, q, sk);
    }
    return hipGetLastError();
}

hipError_t launch_bf16x3(int id, const Problem& p, hipStream_t s, const Arena* arena) {
    if (id < 0 || id >= bf16x3_config_count()) return hipErrorInvalidValue;
    switch (id) {
        case 6: return launch_x3<X3_128x128, true>(p, s, arena);
        case 7: return launch_x3<X3_64x64, true>(p, s, arena);
        case 9: return launch_x3_tail<X3_256x256_w8>(p, s, arena);
        case 10: return launch_x3<X3_128x128_2cu, true, kX3WideSplitKtiles>(p, s, arena);
        case 11: return launch_x3_streamk<X3_256x256_w8>(p, s, arena);
        case 12: return launch_x3_packed(p, s, arena, false);
        case 13: return launch_x3_packed(p, s, arena, true);
        default: break;
    }
    hipError_t e = hipErrorInvalidValue;
    dispatch_x3(id, [&](auto tag) {
        e = launch_x3<typename decltype(tag)::type, false>(p, s, arena);
    });
    return e;
}

}  // namespace sgemm_amd
