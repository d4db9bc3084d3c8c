// common.hpp -- shared device/host definitions for the gfx950 SGEMM kernels.
//
// Column-major throughout: element (i, j) of an (rows x cols) operand with
// leading dimension ld lives at ptr[i + j*ld].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgemm_amd {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// One GEMM problem as the kernels see it (device-side descriptor).  The same
// struct is the element of the grouped kernel's descriptor array.
struct Problem {
    const float* A;
    const float* B;
    float* C;
    int m, n, k;
    int lda, ldb, ldc;
    float alpha, beta;
    int tiles_m, tiles_n;  // tile grid of this problem for the chosen config
    int tile_begin;        // first global tile index (grouped schedules)
    int cfg;               // config id the tiles were planned with
    int flags;             // kProblemVec4: every 16-byte vector access is aligned
    int reserved;
};

constexpr int kProblemVec4 = 1;

static_assert(sizeof(Problem) % 8 == 0, "descriptor packing");

__host__ __device__ constexpr int div_up(int a, int b) { return (a + b - 1) / b; }

// Number of XCDs the dispatcher round-robins workgroups over (MI355X: 8).
constexpr int kNumXcd = 8;

// Workgroup i is placed on XCD (i mod 8).  Remap the hardware id so that each
// XCD receives a contiguous range of logical tiles: neighbouring tiles share
// A/B panels and should share that XCD's L2.  Valid for any grid size: ids in
// the final partial round keep their identity order.
__device__ inline int xcd_remap(int hw_id, int n_ids) {
    const int full = (n_ids / kNumXcd) * kNumXcd;
    if (hw_id >= full) return hw_id;
    const int per_xcd = full / kNumXcd;
    return (hw_id % kNumXcd) * per_xcd + hw_id / kNumXcd;
}

// Logical tile id -> (tile_m, tile_n), walking the tile grid in column bands of
// `group_n` tile-columns so consecutive ids reuse the same B panels and a small
// set of A panels (the classic grouped ordering, transposed for column-major C).
__device__ inline void tile_coords(int id, int tiles_m, int tiles_n, int group_n, int& tm,
                                   int& tn) {
    const int per_band = group_n * tiles_m;
    const int band = id / per_band;
    const int first_n = band * group_n;
    const int width = min(tiles_n - first_n, group_n);
    const int local = id - band * per_band;
    tm = local / width;
    tn = first_n + local % width;
}

}  // namespace sgemm_amd
