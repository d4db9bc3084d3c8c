// sgemm_bf16x3.hip -- single-problem launches of the BF16x3 tile (tile_x3.hpp).
#include "tile_f32.hpp"
#include "tile_x3.hpp"

namespace sgemm_amd {

constexpr int kX3GroupN = 8;

// WPE: waves per SIMD the register allocation must allow (T::WPE).
// SPLIT: the grid is tiles x splits, split index outermost (as sgemm_f32.hip).
template <class T, bool VEC, bool SPLIT>
__global__ __launch_bounds__(T::NT) __attribute__((amdgpu_waves_per_eu(T::WPE)))
void k_sgemm_bf16x3(Problem P, SplitK sk) {
    __shared__ __attribute__((aligned(16))) char smem[T::LDS_BYTES];
    const int id = xcd_remap(blockIdx.x, gridDim.x);
    const int tiles = P.tiles_m * P.tiles_n;
    const int split = SPLIT ? id / tiles : 0;
    int tm, tn;
    tile_coords(SPLIT ? id - split * tiles : id, P.tiles_m, P.tiles_n, kX3GroupN, tm, tn);
    X3TileRunner<T, VEC>::run(P, tm, tn, smem, SPLIT ? &sk : nullptr, split);
}

// Ids 6-7 run the tiles of ids 0-1 with split-K (splitk = 1); id 8 is a
// 128x128 tile with one accumulator set, two workgroups per CU.
static const ConfigInfo kBf16x3Configs[] = {
    {128, 128, 32, 2, 2, "bf16x3_128x128x32_w2x2_m16_dual"},
    {64, 64, 64, 2, 2, "bf16x3_64x64x64_w2x2_m16_dual"},
    {256, 256, 32, 2, 4, "bf16x3_256x256x32_w2x4_m32_single"},
    {256, 128, 32, 2, 2, "bf16x3_256x128x32_w2x2_m32_single"},
    {128, 128, 32, 2, 2, "bf16x3_128x128x32_w2x2_m32_dual"},
    {256, 128, 16, 2, 2, "bf16x3_256x128x16_w2x2_m32_single_mid"},
    {128, 128, 32, 2, 2, "bf16x3_128x128x32_w2x2_m16_dual_splitk", 1},
    {64, 64, 64, 2, 2, "bf16x3_64x64x64_w2x2_m16_dual_splitk", 1},
    {128, 128, 32, 2, 2, "bf16x3_128x128x32_w2x2_m32_single_2cu"},
};

int bf16x3_config_count() { return (int)(sizeof(kBf16x3Configs) / sizeof(kBf16x3Configs[0])); }
const ConfigInfo& bf16x3_config(int id) { return kBf16x3Configs[id]; }

template <class T, bool ALLOW_SPLIT>
static hipError_t launch_x3(const Problem& p, hipStream_t s) {
    Problem q = p;
    q.tiles_m = div_up(p.m, T::BM);
    q.tiles_n = div_up(p.n, T::BN);
    const long long tiles = (long long)q.tiles_m * q.tiles_n;
    if (tiles <= 0) return hipSuccess;
    if (tiles > 0x7fffffffLL) return hipErrorInvalidValue;
    SplitK sk{nullptr, nullptr, 1, p.k};
    if (ALLOW_SPLIT) {
        constexpr size_t acc_bytes = T::MF == 16 ? 16 : 64;
        hipError_t e = splitk_plan(s, tiles, p.k, T::BK, T::MB * T::NB * acc_bytes * T::NT, &sk);
        if (e != hipSuccess) return e;
    }
    const bool split = ALLOW_SPLIT && sk.splits > 1;
    constexpr bool S = ALLOW_SPLIT;
    const dim3 grid((unsigned)(tiles * sk.splits));
    (void)hipGetLastError();
    if (vec4_ok(p)) {
        if (split) hipLaunchKernelGGL((k_sgemm_bf16x3<T, true, S>), grid, dim3(T::NT), 0, s, q, sk);
        else hipLaunchKernelGGL((k_sgemm_bf16x3<T, true, false>), grid, dim3(T::NT), 0, s, q, sk);
    } else {
        if (split) hipLaunchKernelGGL((k_sgemm_bf16x3<T, false, S>), grid, dim3(T::NT), 0, s, q, sk);
        else hipLaunchKernelGGL((k_sgemm_bf16x3<T, false, false>), grid, dim3(T::NT), 0, s, q, sk);
    }
    return hipGetLastError();
}

hipError_t launch_bf16x3(int id, const Problem& p, hipStream_t s) {
    if (id < 0 || id >= bf16x3_config_count()) return hipErrorInvalidValue;
    switch (id) {
        case 6: return launch_x3<X3_128x128, true>(p, s);
        case 7: return launch_x3<X3_64x64, true>(p, s);
        default: break;
    }
    hipError_t e = hipErrorInvalidValue;
    dispatch_x3(id, [&](auto tag) { e = launch_x3<typename decltype(tag)::type, false>(p, s); });
    return e;
}

}  // namespace sgemm_amd
