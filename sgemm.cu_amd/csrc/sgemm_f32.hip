// sgemm_f32.hip -- single-problem launches of the plain FP32 tile
// (tile_f32.hpp), the naive reference-order kernel and the beta-scale kernel.
#include "tile_f32.hpp"

namespace sgemm_amd {

// Column bands of this many tile-columns in the launch order.
constexpr int kGroupN = 8;

// SPLIT: workgroups [0, sk.dp_tiles) compute tiles [0, sk.dp_tiles) whole;
// the rest are (tile, split) units of the remaining tiles, with the split
// index outermost so that workgroups running together share k-chunks.  The
// dispatcher hands out workgroups in id order, so the split units start as
// the whole tiles finish: a partial last round of whole tiles becomes a full
// round of shorter units.
template <class T, bool SPLIT>
__global__ __launch_bounds__(T::NT) __attribute__((amdgpu_waves_per_eu(T::WPE)))
void k_sgemm_f32(Problem P, SplitK sk) {
    static_assert(!SPLIT || T::KG == 1, "k-group tiles run whole tiles only");
    __shared__ __attribute__((aligned(16))) float smem[T::LDS_FLOATS];
    int tm, tn;
    const int b = blockIdx.x;
    if (!SPLIT || b < sk.dp_tiles) {
        tile_coords(xcd_remap(b, SPLIT ? sk.dp_tiles : (int)gridDim.x), P.tiles_m, P.tiles_n,
                    kGroupN, tm, tn);
        F32TileRunner<T>::run(P, tm, tn, smem);
    } else {
        const int units = (int)gridDim.x - sk.dp_tiles;
        const int u = xcd_remap(b - sk.dp_tiles, units);
        const int count = units / sk.splits;  // split tiles in this launch
        const int split = u / count;
        const int slot = u - split * count;
        tile_coords(sk.dp_tiles + slot, P.tiles_m, P.tiles_n, kGroupN, tm, tn);
        F32TileRunner<T>::run(P, tm, tn, smem, &sk, split, slot);
    }
}

// Stream-K over the tail, as k_sgemm_bf16x3_sk: workgroups [0, dp_tiles)
// compute the tiles of the full rounds whole; workgroup dp_tiles + g of the
// rest runs k-tile iterations [g * iters, (g + 1) * iters) of the remaining
// tiles, tile-major with k-tiles in order; a tile cut at a workgroup boundary
// is reduced from its pieces in k order by the last of them to finish.
template <class T>
__global__ __launch_bounds__(T::NT) __attribute__((amdgpu_waves_per_eu(T::WPE)))
void k_sgemm_f32_sk(Problem P, SplitK sk) {
    __shared__ __attribute__((aligned(16))) float smem[T::LDS_FLOATS];
    int tm, tn;
    const int b = blockIdx.x;
    if (b < sk.dp_tiles) {
        tile_coords(xcd_remap(b, sk.dp_tiles), P.tiles_m, P.tiles_n, kGroupN, tm, tn);
        F32TileRunner<T>::run(P, tm, tn, smem);
        return;
    }
    streamk_pieces(P, sk, b, T::BK, [&](int tile, const KPiece& pc) {
        tile_coords(tile, P.tiles_m, P.tiles_n, kGroupN, tm, tn);
        F32TileRunner<T>::run_piece(P, tm, tn, smem, &sk, pc);
    });
}

// Ids 6-8 run the tiles of ids 0-2 with split-K (splitk = 1).  Id 9 runs
// the tile of id 5 with a tail split (splitk = 2): whole tiles for every full
// round of the resident workgroups, split-K for the tiles of the last,
// partial round.  Id 10 runs the tile of id 5 with split-K over the whole
// grid, each split at least eight k-tiles deep.  Id 11 runs the tile of id
// 5 with stream-K over the tiles of a partial last round (splitk = 3).  Id
// 12 is a 64x64x64 tile with two k groups of four waves (F32_64x64_kg2).
// Id 15 is the 256x256x32 tile of f32_256.hip (one wave per SIMD), with id
// 9 standing in where its preconditions do not hold; id 16 is id 15 with
// stream-K over the tiles of a partial last round.
static const ConfigInfo kF32Configs[] = {
    {128, 128, 32, 2, 2, "f32_128x128x32_w2x2"},
    {64, 64, 32, 2, 2, "f32_64x64x32_w2x2"},
    {128, 64, 32, 2, 2, "f32_128x64x32_w2x2"},
    {256, 128, 16, 2, 2, "f32_256x128x16_w2x2"},
    {128, 128, 32, 2, 2, "f32_128x128x32_w2x2_akmajor"},
    {128, 128, 16, 2, 2, "f32_128x128x16_w2x2_3wpe"},
    {128, 128, 32, 2, 2, "f32_128x128x32_w2x2_splitk", 1},
    {64, 64, 32, 2, 2, "f32_64x64x32_w2x2_splitk", 1},
    {128, 64, 32, 2, 2, "f32_128x64x32_w2x2_splitk", 1},
    {128, 128, 16, 2, 2, "f32_128x128x16_w2x2_3wpe_tailsplit", 2},
    {128, 128, 16, 2, 2, "f32_128x128x16_w2x2_3wpe_splitk", 1},
    {128, 128, 16, 2, 2, "f32_128x128x16_w2x2_3wpe_streamk", 3},
    {64, 64, 64, 2, 2, "f32_64x64x64_w2x2_kg2"},
    {128, 128, 32, 2, 2, "f32_128x128x32_w2x2_cprefetch"},
    {128, 64, 16, 2, 2, "f32_128x64x16_w2x2_cprefetch"},
    {256, 256, 32, 2, 2, "f32_256x256x32_w2x2_1wpe"},
    {256, 256, 32, 2, 2, "f32_256x256x32_w2x2_1wpe_streamk", 3},
};

int f32_config_count() { return (int)(sizeof(kF32Configs) / sizeof(kF32Configs[0])); }
const ConfigInfo& f32_config(int id) { return kF32Configs[id]; }

template <class T, bool ALLOW_SPLIT, int MIN_KTILES = 4>
static hipError_t launch_tile(const Problem& p, hipStream_t s, const Arena* arena) {
    Problem q = p;
    q.tiles_m = div_up(p.m, T::BM);
    q.tiles_n = div_up(p.n, T::BN);
    q.tile_begin = 0;
    const long long tiles = (long long)q.tiles_m * q.tiles_n;
    if (tiles <= 0) return hipSuccess;
    if (tiles > 0x7fffffffLL) return hipErrorInvalidValue;
    SplitK sk{nullptr, nullptr, 1, p.k, 0};
    if (ALLOW_SPLIT) {
        hipError_t e =
            splitk_plan(s, tiles, p.k, T::BK, (size_t)T::MB * T::NB * T::NT * 16, &sk, true,
                        MIN_KTILES, arena);
        if (e != hipSuccess) return e;
    }
    const bool split = ALLOW_SPLIT && sk.splits > 1;
    const dim3 grid((unsigned)(tiles * sk.splits));
    (void)hipGetLastError();  // the status read below must belong to this launch
    constexpr bool S = ALLOW_SPLIT;
    if (split) hipLaunchKernelGGL((k_sgemm_f32<T, S>), grid, dim3(T::NT), 0, s, q, sk);
    else hipLaunchKernelGGL((k_sgemm_f32<T, false>), grid, dim3(T::NT), 0, s, q, sk);
    return hipGetLastError();
}

// Workgroups of the split (SK: stream-K) kernel of tile T resident on the
// device at once.
template <class T, bool SK = false>
static long long resident_workgroups() {
    static const long long n = [] {
        int per_cu = 0;
        void (*kernel)(Problem, SplitK) = nullptr;
        if constexpr (SK) kernel = k_sgemm_f32_sk<T>;
        else kernel = k_sgemm_f32<T, true>;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, T::NT, 0) != hipSuccess ||
            per_cu < 1) {
            (void)hipGetLastError();
            per_cu = 1;
        }
        return (long long)per_cu * device_cu_count();
    }();
    return n;
}

// Tail split: the tiles of full rounds run whole; the r tiles of a partial
// last round are each cut into the number of k-ranges tail_split_count
// picks.  Falls back to one launch of whole tiles when there is no partial
// round, when the split would not pay, or when the stream is being captured
// and its workspace would have to grow.
template <class T>
static hipError_t launch_tail_split(const Problem& p, hipStream_t s, const Arena* arena) {
    Problem q = p;
    q.tiles_m = div_up(p.m, T::BM);
    q.tiles_n = div_up(p.n, T::BN);
    q.tile_begin = 0;
    const long long tiles = (long long)q.tiles_m * q.tiles_n;
    if (tiles <= 0) return hipSuccess;
    if (tiles > 0x7fffffffLL) return hipErrorInvalidValue;
    const long long slots = resident_workgroups<T>();
    const long long rem = tiles % slots;
    const size_t split_bytes = (size_t)T::MB * T::NB * T::NT * 16;
    const int per_cu = (int)(slots / device_cu_count());
    double tail_cost = 1.0;
    const int splits =
        tail_split_count(tiles, slots, per_cu, p.k, T::BK, split_bytes,
                         per_cu * 2.0 * T::BM * T::BN * p.k / kF32FlopsPerCuUs, &tail_cost);
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
    (void)hipGetLastError();  // the status read below must belong to this launch
    if (sk.splits > 1) {
        sk.dp_tiles = (int)(tiles - rem);
        const dim3 grid((unsigned)(tiles - rem + rem * sk.splits));
        hipLaunchKernelGGL((k_sgemm_f32<T, true>), grid, dim3(T::NT), 0, s, q, sk);
    } else {
        const dim3 grid((unsigned)tiles);
        hipLaunchKernelGGL((k_sgemm_f32<T, false>), grid, dim3(T::NT), 0, s, q, sk);
    }
    return hipGetLastError();
}

// Stream-K over the tail (launch_x3_streamk's scheme on config 5's tile):
// the r = tiles % slots tiles after the full rounds (all tiles without a full
// round) spread, k-tile by k-tile, over up to `slots` workgroups of w k-tiles
// each (streamk_iters).  Falls back to one launch of whole tiles when there
// is no partial round, or when the stream is being captured and its
// workspace would have to grow.
template <class T>
static hipError_t launch_streamk(const Problem& p, hipStream_t s, const Arena* arena) {
    Problem q = p;
    q.tiles_m = div_up(p.m, T::BM);
    q.tiles_n = div_up(p.n, T::BN);
    q.tile_begin = 0;
    const long long tiles = (long long)q.tiles_m * q.tiles_n;
    if (tiles <= 0) return hipSuccess;
    if (tiles > 0x7fffffffLL) return hipErrorInvalidValue;
    const long long slots = resident_workgroups<T, true>();
    const long long rem = tiles % slots;
    const size_t split_bytes = (size_t)T::MB * T::NB * T::NT * 16;
    SplitK sk{nullptr, nullptr, 1, p.k, 0};
    long long nwg = 0;
    const hipError_t e =
        streamk_plan(s, tiles, slots, p.k, T::BK, split_bytes, kF32StreamKPieceIters, &sk, &nwg,
                     arena);
    if (e != hipSuccess) return e;
    (void)hipGetLastError();
    if (nwg > 0) {
        sk.dp_tiles = (int)(tiles - rem);
        const dim3 grid((unsigned)(tiles - rem + nwg));
        hipLaunchKernelGGL((k_sgemm_f32_sk<T>), grid, dim3(T::NT), 0, s, q, sk);
    } else {
        const dim3 grid((unsigned)tiles);
        hipLaunchKernelGGL((k_sgemm_f32<T, false>), grid, dim3(T::NT), 0, s, q, sk);
    }
    return hipGetLastError();
}

hipError_t launch_f32(int id, const Problem& p, hipStream_t s, const Arena* arena) {
    if (id < 0 || id >= f32_config_count()) return hipErrorInvalidValue;
    const bool dl = columns_16b_aligned(p);
    switch (id) {
        case 5:
            return dl ? launch_tile<F32_128x128_k16_dl, false>(p, s, arena)
                      : launch_tile<F32_128x128_k16, false>(p, s, arena);
        case 6: return launch_tile<F32_128x128, true>(p, s, arena);
        case 1:
            return dl ? launch_tile<F32_64x64_dl, false>(p, s, arena)
                      : launch_tile<F32_64x64, false>(p, s, arena);
        case 7:
            return dl ? launch_tile<F32_64x64_dl, true>(p, s, arena)
                      : launch_tile<F32_64x64, true>(p, s, arena);
        case 8: return launch_tile<F32_128x64, true>(p, s, arena);
        case 9:
            return dl ? launch_tail_split<F32_128x128_k16_dl>(p, s, arena)
                      : launch_tail_split<F32_128x128_k16>(p, s, arena);
        case 10:
            return dl ? launch_tile<F32_128x128_k16_dl, true, 8>(p, s, arena)
                      : launch_tile<F32_128x128_k16, true, 8>(p, s, arena);
        // Stream-K keeps the register-staged tile: with the direct-loaded
        // one (at three workgroups per CU, as its piece loop needs more than
        // 128 VGPRs) it ran 118 against 130 TFLOP/s at 3584^3.
        case 11: return launch_streamk<F32_128x128_k16>(p, s, arena);
        case 12: return launch_tile<F32_64x64_kg2, false>(p, s, arena);
        case 15:
        case 16:
            if (f32_256_supported(p)) return launch_f32_256(p, s, id == 16, arena);
            return dl ? launch_tail_split<F32_128x128_k16_dl>(p, s, arena)
                      : launch_tail_split<F32_128x128_k16>(p, s, arena);
        default: break;
    }
    hipError_t e = hipErrorInvalidValue;
    dispatch_f32(id, [&](auto tag) {
        e = launch_tile<typename decltype(tag)::type, false>(p, s, arena);
    });
    return e;
}

// ------------------------------------------------------------ naive kernel
// Column-major restatement of kernel_basic (src/kernels/basic.cuh:15-25): one
// thread per output, sequential k with FMA, *alpha, then fma(C, beta, acc).
// C is always read, as in the reference.  WIDE: the fallback for leading
// dimensions the tiles' 32-bit buffer offsets cannot reach -- 64-bit
// addressing, FP32 arithmetic in either form, and BLAS beta semantics
// (beta == 0 stores without reading C).
template <bool WIDE>
__global__ __launch_bounds__(256) void k_sgemm_naive(Problem P) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= P.m || j >= P.n) return;
    const float* a = P.A + i;
    const float* b = P.B + (size_t)j * P.ldb;
    float acc = 0.f;
    for (int p = 0; p < P.k; ++p) acc = __builtin_fmaf(a[(size_t)p * P.lda], b[p], acc);
    acc *= P.alpha;
    float* c = P.C + i + (size_t)j * P.ldc;
    *c = (WIDE && P.beta == 0.f) ? acc : __builtin_fmaf(*c, P.beta, acc);
}

template <bool WIDE>
static hipError_t launch_naive(const Problem& p, hipStream_t s) {
    if (p.m <= 0 || p.n <= 0) return hipSuccess;
    (void)hipGetLastError();
    // Chunk the column range to stay within the grid's y limit.
    for (long long j0 = 0; j0 < p.n; j0 += 4LL * 65535) {
        Problem q = p;
        q.n = (int)std::min<long long>(p.n - j0, 4LL * 65535);
        q.B = p.B + (size_t)j0 * p.ldb;
        q.C = p.C + (size_t)j0 * p.ldc;
        hipLaunchKernelGGL(k_sgemm_naive<WIDE>, dim3(div_up(q.m, 64), div_up(q.n, 4)), dim3(256), 0,
                           s, q);
    }
    return hipGetLastError();
}

hipError_t launch_basic(const Problem& p, hipStream_t s) { return launch_naive<false>(p, s); }
hipError_t launch_wide(const Problem& p, hipStream_t s) { return launch_naive<true>(p, s); }

// --------------------------------------------------------- C := beta * C
// Used when k == 0 or alpha == 0 (no product term).  beta == 0 stores zeros
// without reading C.
__global__ __launch_bounds__(256) void k_scale_c(Problem P) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= P.m) return;
    float* c = P.C + i + (size_t)j * P.ldc;
    *c = P.beta != 0.f ? P.beta * *c : 0.f;
}

hipError_t launch_scale_c(const Problem& p, hipStream_t s) {
    if (p.m <= 0 || p.n <= 0) return hipSuccess;
    (void)hipGetLastError();
    if (p.n > 65535) {
        // Chunk the column range to stay within the grid's y limit.
        for (int j0 = 0; j0 < p.n; j0 += 65535) {
            Problem q = p;
            q.C = p.C + (size_t)j0 * p.ldc;
            q.n = p.n - j0 < 65535 ? p.n - j0 : 65535;
            hipLaunchKernelGGL(k_scale_c, dim3(div_up(q.m, 256), q.n), dim3(256), 0, s, q);
        }
    } else {
        hipLaunchKernelGGL(k_scale_c, dim3(div_up(p.m, 256), p.n), dim3(256), 0, s, p);
    }
    return hipGetLastError();
}

}  // namespace sgemm_amd
