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
    int flags;             // reserved, zero
    int reserved;
};

// Split-K state of one workgroup: it computes k in [split*kchunk, ...) of
// tile `tile`, dumps its accumulators to `partials`, and the last of the
// `splits` workgroups of that tile to arrive (counted in counters[tile]) sums
// all partials in split order 0, 1, ..., splits-1 -- so the result does not
// depend on which workgroup finishes last -- and runs the epilogue.
//
// Coherence across XCDs (each has its own L2): the partials are written and
// read with the sc1 cache policy, the policy of agent-scope atomic loads and
// stores, so they pass through to the device-wide coherence point instead of
// sitting in one XCD's L2.  The writer waits for its stores to complete before
// the counter increment, which is then all the release the reader needs; no
// agent-scope fence (a whole-L2 writeback plus invalidate) is issued.
struct SplitK {
    float* partials;  // [tiles][splits][accumulators][threads] float4s
    int* counters;    // [tiles], zero between launches
    int splits;
    int kchunk;       // multiple of the tile's BK
    // Tail split: workgroups [0, dp_tiles) compute whole tiles [0, dp_tiles)
    // unsplit; the workgroups after them split the remaining tiles' k
    // ranges.  0: every workgroup is a split unit.
    int dp_tiles;
    // Stream-K over the tiles after dp_tiles (iters > 0): workgroup g of the
    // stream-K part runs k-tile iterations [g * iters, (g + 1) * iters) of
    // those tiles' iteration space (tile-major, k-tiles in order), so a tile
    // is cut into pieces at workgroup boundaries.  A cut tile's partials and
    // counter sit at the index of the workgroup holding its last k-tile,
    // `pieces` partial slots apart.
    int iters = 0;
    int pieces = 0;
};

// The k range one workgroup accumulates for a tile: k in [kbeg, kend), piece
// q of the tile's n pieces (in k order).  n > 1: the partials are reduced
// through workspace region `region` of `stride` piece slots, and counter
// `region`.
struct KPiece {
    int kbeg, kend, q, n, region, stride;
};

// Cache policy of the partials' loads and stores: sc1 (agent coherence).
constexpr int kPolicyAgent = 0x10;

__device__ inline void store_agent(__amdgpu_buffer_rsrc_t r, int off, f32x4 v) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, off, 0, kPolicyAgent);
}
__device__ inline f32x4 load_agent(__amdgpu_buffer_rsrc_t r, int off) {
    return __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, kPolicyAgent));
}

// Called by every thread of the workgroup after its partial stores.  Returns
// true in the workgroup that arrives last of the n for `tile` (and resets the
// counter for the next launch).
__device__ inline bool splitk_arrive(const SplitK& sk, int tile, int* flag_lds, int n) {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);  // keep the stores above the wait
    __builtin_amdgcn_s_waitcnt(0);            // this wave's partial stores completed
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();                          // ... and every other wave's
    if (threadIdx.x == 0)
        *flag_lds = __hip_atomic_fetch_add(&sk.counters[tile], 1, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const bool last = *flag_lds == n - 1;
    if (last && threadIdx.x == 0)
        __hip_atomic_store(&sk.counters[tile], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return last;
}

static_assert(sizeof(Problem) % 8 == 0, "descriptor packing");

__host__ __device__ constexpr int div_up(int a, int b) { return (a + b - 1) / b; }

// Elements a column-major (rows x cols, ld) operand spans.
__host__ __device__ inline long long extent(int rows, int cols, int ld) {
    return (rows <= 0 || cols <= 0) ? 0 : (long long)ld * (cols - 1) + rows;
}

// Leading dimensions above this would overflow a tile's 32-bit buffer
// offsets (up to 256 columns of B per tile); the tiled kernels reject them.
constexpr int kMaxLd = 1 << 21;

// Buffer resource over `bytes` bytes from `base` (gfx9-family word 3).  Loads
// through it at offsets >= bytes return 0 and never fault, so edge tiles need
// no branches around their loads.
__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void* base, long long bytes) {
    const unsigned n = bytes <= 0 ? 0u : bytes >= 0xffffffffLL ? 0xffffffffu : (unsigned)bytes;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)n,
                                             0x00020000);
}

// Four consecutive floats at byte offset `off` of a buffer: one 16-byte load.
// Floats are 4-byte aligned and a multi-dword access needs only dword
// alignment, so this serves every leading dimension; the range check is per
// dword, so a load that reaches past the end of the buffer still returns the
// floats below it (and 0 beyond).
__device__ inline f32x4 bload4(__amdgpu_buffer_rsrc_t r, int off) {
    return __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}

// 16 bytes per lane from byte offset `off` of buffer `r` into LDS: lane i's
// data lands at lds + 4 * i floats (buffer_load_dwordx4 with the lds bit;
// `lds` must be the same for every lane).  Out-of-range dwords write 0.
// Completion is counted by vmcnt, not lgkmcnt: see wait_vmem.
// soff: a wave-uniform byte offset added in the instruction's scalar offset.
__device__ inline void buffer_load_lds16(__amdgpu_buffer_rsrc_t r, float* lds, int off,
                                         int soff = 0) {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(
        r, (__attribute__((address_space(3))) void*)lds, 16, off, soff, 0, 0);
}

// s_waitcnt vmcnt(0): every vector memory load of this wave has completed,
// including the LDS writes of buffer_load_lds16.
__device__ inline void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Rows r..r+3 of one column, zero outside [0, valid): one 16-byte load when
// all four are valid (dword alignment suffices, see bload4).
__device__ inline f32x4 load4(const float* p, int valid) {
    f32x4 v = {0.f, 0.f, 0.f, 0.f};
    if (valid >= 4) {
        __builtin_memcpy(&v, p, sizeof(v));
        return v;
    }
    if (valid > 0) v[0] = p[0];
    if (valid > 1) v[1] = p[1];
    if (valid > 2) v[2] = p[2];
    return v;
}

// The first valid (<= 4) of rows r..r+3 of one column.
__device__ inline void store4(float* p, f32x4 v, int valid) {
    if (valid >= 4) {
        __builtin_memcpy(p, &v, sizeof(v));
        return;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
        if (r < valid) p[r] = v[r];
}

// Floats at the start of a tile's LDS that the epilogue leaves alone: the
// split-K arrival flag lives there while the other waves still read it.
constexpr int kEpilogueLdsSkip = 16;

// LDS floats the epilogue of a tile of `waves` waves with TM-row wave tiles
// and MF-column MFMA blocks needs (kEpilogueLdsSkip included).
constexpr int epilogue_lds_floats(int waves, int TM, int MF) {
    return kEpilogueLdsSkip + waves * MF * (TM + 4);
}

// Epilogue of one MF-column block of a wave's TM-row tile, through LDS.  The
// MFMA output layout gives a lane four consecutive rows of one column per
// register group, so a store straight from it touches MF columns with 16 to
// 64 bytes each -- part of a cache line per column.  Here the block goes to a
// wave-private LDS area [column][row] (row stride TM + 4 floats) and back so
// that a lane holds four consecutive rows of a column and the wave's 64
// lanes cover 256 / TM whole columns of the tile: every load and store of C
// is a run of whole 16-byte pieces down a column.  The arithmetic per
// element is unchanged, fma(beta, c, alpha * acc) or alpha * acc with beta ==
// 0 (C not read).
//   wlds:  MF * (TM + 4) floats of LDS used by this wave only; the LDS must
//          be free (every wave of the workgroup past its last use of it).
//   m0w:   the wave tile's first row; col0: the block's first column.
//   vals(i, q): the block's MFMA block i, register group q (4 rows).
//   c_at(p, cc, valid): the C values of read-back pass p at cc (beta != 0).
// DS instructions of one wave complete in order, so the read-back sees the
// writes without a barrier.
template <int TM, int MF, int MB, class Vals, class CAt>
__device__ inline void epilogue_block(float* wlds, const Problem& P, int m0w, int col0,
                                      Vals&& vals, CAt&& c_at) {
    constexpr int TMP = TM + 4;
    constexpr int GROUPS = MF == 16 ? 1 : 4;
    constexpr int LPC = TM / 4;   // lanes per column on read-back
    constexpr int CPP = 64 / LPC; // columns per read-back pass
    static_assert(MF == 16 || MF == 32, "MFMA block");
    static_assert(64 % LPC == 0 && MF % CPP == 0, "read-back mapping");
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int q = 0; q < GROUPS; ++q) {
            const int row = MF * i + (MF == 16 ? 4 * (lane / 16) : 8 * q + 4 * (lane / 32));
            *reinterpret_cast<f32x4*>(&wlds[(lane % MF) * TMP + row]) = vals(i, q);
        }
    __builtin_amdgcn_wave_barrier();
    const int c_in = lane / LPC, r = 4 * (lane % LPC);
    const int row = m0w + r;
    const int valid = min(max(P.m - row, 0), 4);
    const float alpha = P.alpha, beta = P.beta;
#pragma unroll
    for (int p = 0; p < MF / CPP; ++p) {
        const int c = p * CPP + c_in;
        const f32x4 a = *reinterpret_cast<const f32x4*>(&wlds[c * TMP + r]);
        const int col = col0 + c;
        if (col >= P.n || valid == 0) continue;
        float* cc = P.C + (size_t)col * P.ldc + row;
        f32x4 v;
        if (beta != 0.f) {
            const f32x4 c0 = c_at(p, cc, valid);
#pragma unroll
            for (int e = 0; e < 4; ++e) v[e] = __builtin_fmaf(beta, c0[e], alpha * a[e]);
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) v[e] = alpha * a[e];
        }
        store4(cc, v, valid);
    }
    __builtin_amdgcn_wave_barrier();
}

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

// Stream-K part of a launch whose first sk.dp_tiles workgroups run whole
// tiles: calls piece(tile, KPiece) for each piece of stream-K workgroup
// `b - sk.dp_tiles`, in order.  That workgroup runs k-tile iterations [g *
// iters, (g + 1) * iters) of the tiles after dp_tiles, tile-major with bk-deep
// k-tiles in order (g = the XCD-remapped index); a tile's piece q of n is
// reduced through region `last` (the workgroup holding its last k-tile).
// Every thread of the workgroup calls this; `piece` must leave the workgroup
// converged, and the barrier between pieces keeps the next piece's LDS
// stores behind every wave's read of the previous one's arrival flag.
template <class F>
__device__ inline void streamk_pieces(const Problem& P, const SplitK& sk, int b, int bk,
                                      F&& piece) {
    const int nwg = (int)gridDim.x - sk.dp_tiles;
    const long long g = xcd_remap(b - sk.dp_tiles, nwg);
    const int I = div_up(P.k, bk);  // k-tiles per tile
    const long long L = ((long long)P.tiles_m * P.tiles_n - sk.dp_tiles) * I;
    const long long w = sk.iters;
    long long x = g * w;
    const long long end = min(L, x + w);
    while (x < end) {
        const int t = (int)(x / I);
        const int kt0 = (int)(x - (long long)t * I);
        const int kt1 = (int)min((long long)I, kt0 + (end - x));
        const int first = (int)((long long)t * I / w);          // workgroup of k-tile 0
        const int last = (int)(((long long)t * I + I - 1) / w);  // ... and of the last
        piece(sk.dp_tiles + t, KPiece{kt0 * bk, min(P.k, kt1 * bk), (int)g - first,
                                      last - first + 1, last, sk.pieces});
        x = (long long)t * I + kt1;
        __syncthreads();
    }
}

}  // namespace sgemm_amd
