// sgemm_f32.hip -- single-problem launches of the plain FP32 tile
// (tile_f32.hpp), the naive reference-order kernel and the beta-scale kernel.
#include "tile_f32.hpp"

namespace sgemm_amd {

// Column bands of this many tile-columns in the launch order.
constexpr int kGroupN = 8;

// SPLIT: the grid is tiles x splits; workgroup id -> (split, tile) with the
// split index outermost, so workgroups running together share k-chunks.
template <class T, bool VEC, bool SPLIT>
__global__ __launch_bounds__(T::NT) __attribute__((amdgpu_waves_per_eu(T::WPE)))
void k_sgemm_f32(Problem P, SplitK sk) {
    __shared__ __attribute__((aligned(16))) float smem[T::LDS_FLOATS];
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    const int tiles = P.tiles_m * P.tiles_n;
    const int split = SPLIT ? id / tiles : 0;
    int tm, tn;
    tile_coords(SPLIT ? id - split * tiles : id, P.tiles_m, P.tiles_n, kGroupN, tm, tn);
    F32TileRunner<T, VEC>::run(P, tm, tn, smem, SPLIT ? &sk : nullptr, split);
}

// Ids 6-8 run the tiles of ids 0-2 with split-K (splitk = 1).
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
};

int f32_config_count() { return (int)(sizeof(kF32Configs) / sizeof(kF32Configs[0])); }
const ConfigInfo& f32_config(int id) { return kF32Configs[id]; }

template <class T, bool ALLOW_SPLIT>
static hipError_t launch_tile(const Problem& p, hipStream_t s) {
    Problem q = p;
    q.tiles_m = div_up(p.m, T::BM);
    q.tiles_n = div_up(p.n, T::BN);
    const long long tiles = (long long)q.tiles_m * q.tiles_n;
    if (tiles <= 0) return hipSuccess;
    if (tiles > 0x7fffffffLL) return hipErrorInvalidValue;
    SplitK sk{nullptr, nullptr, 1, p.k};
    if (ALLOW_SPLIT) {
        hipError_t e = splitk_plan(s, tiles, p.k, T::BK, (size_t)T::MB * T::NB * T::NT * 16, &sk);
        if (e != hipSuccess) return e;
    }
    const bool split = ALLOW_SPLIT && sk.splits > 1;
    const dim3 grid((unsigned)(tiles * sk.splits));
    (void)hipGetLastError();  // the status read below must belong to this launch
    constexpr bool S = ALLOW_SPLIT;
    if (vec4_ok(p)) {
        if (split) hipLaunchKernelGGL((k_sgemm_f32<T, true, S>), grid, dim3(T::NT), 0, s, q, sk);
        else hipLaunchKernelGGL((k_sgemm_f32<T, true, false>), grid, dim3(T::NT), 0, s, q, sk);
    } else {
        if (split) hipLaunchKernelGGL((k_sgemm_f32<T, false, S>), grid, dim3(T::NT), 0, s, q, sk);
        else hipLaunchKernelGGL((k_sgemm_f32<T, false, false>), grid, dim3(T::NT), 0, s, q, sk);
    }
    return hipGetLastError();
}

hipError_t launch_f32(int id, const Problem& p, hipStream_t s) {
    if (id < 0 || id >= f32_config_count()) return hipErrorInvalidValue;
    switch (id) {
        case 6: return launch_tile<F32_128x128, true>(p, s);
        case 7: return launch_tile<F32_64x64, true>(p, s);
        case 8: return launch_tile<F32_128x64, true>(p, s);
        default: break;
    }
    hipError_t e = hipErrorInvalidValue;
    dispatch_f32(id, [&](auto tag) { e = launch_tile<typename decltype(tag)::type, false>(p, s); });
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
