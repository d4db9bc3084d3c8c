// f32_256.hip -- plain FP32 config 15: a 256x256x32 tile of four waves, each
// wave a 128x128 block of v_mfma_f32_16x16x4_f32 (64 blocks, 256
// accumulators in AGPRs), one wave per SIMD and one workgroup per CU.
//
// With one wave per SIMD nothing else covers a stall, so the k loop is
// pipelined inside each wave:
//   - A and B reach LDS by direct loads (buffer_load_dwordx4 ... lds) into
//     two 64 KB stages; the loads of k-tile t + 2 are issued while k-tile t
//     is multiplied, one per 16 MFMAs, so each has a whole k-tile to land.
//   - The MFMA operands of each 16-k half of a k-tile are read from LDS into
//     registers while the other half is multiplied: registers f0 always hold
//     the first half of a k-tile, f1 the second.  Half 1 of k-tile t is read
//     during half 0; half 0 of k-tile t + 1 during half 1, after the one
//     barrier per k-tile, which also frees stage t for the loads of t + 2.
//   - The LDS reads are spread one per 16 MFMAs by the scheduling barriers.
// A lane's MFMA operands are read four at a time (ds_read_b128, no bank
// conflicts):
//   - A is stored as LDS loads it, [k][m] (m contiguous).  MFMA block i of a
//     wave holds, in its row d, wave-tile row 64 (i / 4) + 4 d + i % 4, so
//     the 16 bytes at row 64 h + 4 d of one k give a lane its operand of the
//     four blocks 4h .. 4h + 3.
//   - B is stored [n][k], 32 k per column, the 16-byte chunks of column n at
//     position chunk ^ ((n / 2) % 8); the 16 bytes of one chunk are one
//     block's operand for four consecutive k-steps.
//   - Lane group g supplies k = kk + 4g + s at MFMA step s of each 16-k
//     half, for A and B alike: the same k sets, in the same order, as config
//     5's tile, so config 15 gives config 5's result bits.
// The epilogue goes through LDS per 16-column block, so every C access is a
// 512-byte run down a column: fma(beta, c, alpha * acc), or alpha * acc with
// beta == 0 (C not read).
//
// Preconditions (f32_256_supported): 16-byte aligned A and B columns, k a
// multiple of 32 and at least 64, byte offsets within a tile's operand
// panels below 2^31.  Partial tiles in m and n are fine: each operand's
// buffer resource is bounded by its extent, so no load faults; rows past m
// and columns past n reach only C entries that are never stored, and the
// epilogue stores only inside C.  Elsewhere config 15 runs config 9.
#include "tile_f32.hpp"

namespace sgemm_amd {

namespace {

constexpr int kBM = 256, kBN = 256, kBK = 32;
constexpr int kAStage = kBK * kBM;          // floats, [k][m]
constexpr int kBStage = kBN * kBK;          // floats, [n][k]
constexpr int kStage = kAStage + kBStage;   // 64 KB
constexpr int kLdsFloats = 2 * kStage;      // 128 KB
constexpr int kEpiStride = 132;             // epilogue LDS rows: 128 + 4
constexpr int kGroupN256 = 8;
static_assert(4 * 16 * kEpiStride <= kLdsFloats, "epilogue LDS");

// One 16-k half of a k-tile's MFMA operands for one lane.
struct Half {
    f32x4 a[2][4];  // [h][s]: component e is block 4h + e's operand at step s
    f32x4 b[8];     // [j]: component s is block j's operand at step s
};

__device__ __forceinline__ void read_half(Half& f, const float* st, int hh, int wm, int wn,
                                          int d, int g) {
    const float* As = st;
    const float* Bs = st + kAStage;
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int s = 0; s < 4; ++s)
            f.a[h][s] = *reinterpret_cast<const f32x4*>(
                &As[(16 * hh + 4 * g + s) * kBM + 128 * wm + 64 * h + 4 * d]);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int n = 128 * wn + 16 * j + d;
        f.b[j] = *reinterpret_cast<const f32x4*>(
            &Bs[n * kBK + 4 * ((4 * hh + g) ^ ((n >> 1) & 7))]);
    }
}

// MFMA steps [S0, S1) of one half, all 64 blocks per step.
template <int S0 = 0, int S1 = 4>
__device__ __forceinline__ void mma_half(f32x4 (&acc)[8][8], const Half& f) {
#pragma unroll
    for (int s = S0; s < S1; ++s)
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int j = 0; j < 8; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(f.a[i >> 2][s][i & 3],
                                                                 f.b[j][s], acc[i][j], 0, 0, 0);
}

// Instruction orders for the scheduler (__builtin_amdgcn_sched_group_barrier
// masks: 0x008 MFMA, 0x020 VMEM read, 0x100 DS read): N groups of (V direct
// loads, R operand reads, sixteen MFMAs).  The compiler keeps every LDS read
// behind the direct loads into LDS issued before it (it cannot tell the two
// stages apart), so the half that loads issues its loads among its first
// 128 MFMAs and its reads among the last 128, in two scheduling regions.
template <int N, int V, int R>
__device__ __forceinline__ void schedule_groups() {
#pragma unroll
    for (int r = 0; r < N; ++r) {
        if constexpr (V > 0) __builtin_amdgcn_sched_group_barrier(0x020, V, 0);
        if constexpr (R > 0) __builtin_amdgcn_sched_group_barrier(0x100, R, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 16, 0);
    }
}

// Accumulates k-tiles [kt0, kt1) of tile (m0, n0) into acc, from zero:
// kt1 - kt0 >= 2, or >= 1 with ONE (stream-K pieces).  Every thread of the
// workgroup calls it; on return the workgroup may still be reading LDS (the
// caller's next LDS use follows a barrier).
template <bool ONE>
__device__ __forceinline__ void f256_accumulate(const Problem& P, int m0, int n0, int kt0,
                                                int kt1, float* smem, f32x4 (&acc)[8][8]) {
    const int mrows = min(kBM, P.m - m0), ncols = min(kBN, P.n - n0);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = wave & 1, wn = wave >> 1;
    const int d = lane & 15, g = lane >> 4;
    const int lda = P.lda, ldb = P.ldb, K = P.k;

    // Direct loads: wave w fills A's k rows 8w .. 8w + 7 (one 1 KB row of
    // 256 m per instruction) and B's columns 64w .. 64w + 63 (eight columns
    // of 32 k per instruction, lane l at position l % 8 of column l / 8,
    // which holds chunk (l % 8) ^ ((n / 2) % 8)).  One buffer per operand
    // for the whole k loop.  B's k-tile offset goes in the instruction's
    // scalar offset (no vector arithmetic): its out-of-range columns are out
    // of range at any k.  A's goes in the vector offset, which the buffer's
    // range check covers: rows past m of the last column lie past A's end
    // and must read as 0, not be fetched.
    const auto ra = make_rsrc(P.A + m0, extent(mrows, K, lda) * 4);
    const auto rb = make_rsrc(P.B + (size_t)n0 * ldb, extent(K, ncols, ldb) * 4);
    int a_off[8], b_off[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) a_off[r] = ((8 * wave + r) * lda + 4 * lane) * 4;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int n = 64 * wave + 8 * r + (lane >> 3);
        b_off[r] = (n * ldb + 4 * ((lane & 7) ^ ((n >> 1) & 7))) * 4;
    }
    auto dma = [&](float* st, int kt) {
        const int ka = kt * kBK * lda * 4, kb = kt * kBK * 4;
#pragma unroll
        for (int r = 0; r < 8; ++r) buffer_load_lds16(ra, st + (8 * wave + r) * kBM, a_off[r] + ka);
#pragma unroll
        for (int r = 0; r < 8; ++r)
            buffer_load_lds16(rb, st + kAStage + (64 * wave + 8 * r) * kBK, b_off[r], kb);
    };

#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    // Stage (t & 1) holds k-tile kt0 + t.
    const int KT = kt1 - kt0;
    Half f0, f1;
    dma(smem, kt0);
    if (!ONE || KT > 1) dma(smem + kStage, kt0 + 1);
    wait_vmem();
    __syncthreads();
    read_half(f0, smem, 0, wm, wn, d, g);
    if constexpr (ONE) {
        if (KT == 1) {
            read_half(f1, smem, 1, wm, wn, d, g);
            mma_half(acc, f0);
            mma_half(acc, f1);
            return;
        }
    }

    // This wave's direct loads of the next k-tile have reached LDS and its
    // operand reads of the current one are done; then every wave's are.
    auto sync = [&]() {
        wait_vmem();
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_sched_barrier(0);
    };
    int t = 0;
    for (; t + 2 < KT; ++t) {
        read_half(f1, smem + (t & 1) * kStage, 1, wm, wn, d, g);
        mma_half(acc, f0);
        schedule_groups<16, 0, 1>();
        __builtin_amdgcn_sched_barrier(0);
        sync();
        dma(smem + (t & 1) * kStage, kt0 + t + 2);
        mma_half<0, 2>(acc, f1);
        schedule_groups<8, 2, 0>();
        __builtin_amdgcn_sched_barrier(0);
        read_half(f0, smem + ((t + 1) & 1) * kStage, 0, wm, wn, d, g);
        mma_half<2, 4>(acc, f1);
        schedule_groups<8, 0, 2>();
        __builtin_amdgcn_sched_barrier(0);
    }
    // k-tile KT - 2: nothing left to load.
    read_half(f1, smem + (t & 1) * kStage, 1, wm, wn, d, g);
    mma_half(acc, f0);
    sync();
    read_half(f0, smem + ((t + 1) & 1) * kStage, 0, wm, wn, d, g);
    mma_half(acc, f1);
    ++t;
    // k-tile KT - 1.
    read_half(f1, smem + (t & 1) * kStage, 1, wm, wn, d, g);
    mma_half(acc, f0);
    mma_half(acc, f1);
}

// Epilogue of tile (m0, n0), per 16-column block j: each wave writes its
// 128 x 16 results to its own LDS area as [column][row] (rows of block 4h +
// e's register group q are rows 64h + 16g + 4q + e, so the four blocks of
// one h give four consecutive rows) and reads back two whole columns per
// pass.  Starts with a barrier: every wave is past its last operand read.
__device__ __forceinline__ void f256_epilogue(const Problem& P, int m0, int n0, float* smem,
                                              const f32x4 (&acc)[8][8]) {
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = wave & 1, wn = wave >> 1;
    const int d = lane & 15, g = lane >> 4;
    float* wlds = smem + wave * (16 * kEpiStride);
    const int c_in = lane >> 5, r4 = 4 * (lane & 31);
    const int row = m0 + 128 * wm + r4;
    const int valid = min(max(P.m - row, 0), 4);
    const float alpha = P.alpha, beta = P.beta;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int colb = n0 + 128 * wn + 16 * j;
        f32x4 cv[8];
        if (beta != 0.f) {
#pragma unroll
            for (int p = 0; p < 8; ++p) {
                const int col = colb + 2 * p + c_in;
                cv[p] = (col < P.n && valid > 0) ? load4(P.C + (size_t)col * P.ldc + row, valid)
                                                 : f32x4{0.f, 0.f, 0.f, 0.f};
            }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const f32x4 v = {acc[4 * h][j][q], acc[4 * h + 1][j][q], acc[4 * h + 2][j][q],
                                 acc[4 * h + 3][j][q]};
                *reinterpret_cast<f32x4*>(&wlds[d * kEpiStride + 64 * h + 16 * g + 4 * q]) = v;
            }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            const f32x4 a = *reinterpret_cast<const f32x4*>(&wlds[(2 * p + c_in) * kEpiStride + r4]);
            const int col = colb + 2 * p + c_in;
            f32x4 v;
            if (beta != 0.f) {
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = __builtin_fmaf(beta, cv[p][e], alpha * a[e]);
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = alpha * a[e];
            }
            if (col < P.n && valid > 0) store4(P.C + (size_t)col * P.ldc + row, v, valid);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// One piece's accumulator (i, j) of thread t at byte 16 t + 4096 (8 i + j):
// the per-thread part is the vector offset, the rest the scalar offset, so
// the offsets take no registers of their own.
__device__ __forceinline__ void store_piece(__amdgpu_buffer_rsrc_t r, int voff, int soff,
                                            f32x4 v) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, voff, soff,
                                           kPolicyAgent);
}
__device__ __forceinline__ f32x4 load_piece(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(f32x4,
                              __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, kPolicyAgent));
}

// Config 15: whole tiles.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_sgemm_f32_256(Problem P) {
    __shared__ __attribute__((aligned(16))) float smem[kLdsFloats];
    int tm, tn;
    tile_coords(xcd_remap(blockIdx.x, gridDim.x), P.tiles_m, P.tiles_n, kGroupN256, tm, tn);
    f32x4 acc[8][8];
    f256_accumulate<false>(P, tm * kBM, tn * kBN, 0, P.k / kBK, smem, acc);
    f256_epilogue(P, tm * kBM, tn * kBN, smem, acc);
}

// Config 16: config 15 with stream-K over the tiles of a partial last round.
// Workgroups [0, sk.dp_tiles) compute whole tiles; the rest run the k-tile
// ranges of streamk_pieces over the remaining tiles.  A cut tile's pieces
// store their accumulators (in lane order, 256 KB per piece) and the last to
// arrive adds them in k order, so the result does not depend on which
// piece finishes last.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_sgemm_f32_256_sk(Problem P, SplitK sk) {
    __shared__ __attribute__((aligned(16))) float smem[kLdsFloats];
    int tm, tn;
    const int b = blockIdx.x;
    if (b < sk.dp_tiles) {
        tile_coords(xcd_remap(b, sk.dp_tiles), P.tiles_m, P.tiles_n, kGroupN256, tm, tn);
        f32x4 acc[8][8];
        f256_accumulate<false>(P, tm * kBM, tn * kBN, 0, P.k / kBK, smem, acc);
        f256_epilogue(P, tm * kBM, tn * kBN, smem, acc);
        return;
    }
    streamk_pieces(P, sk, b, kBK, [&](int tile, const KPiece& pc) {
        tile_coords(tile, P.tiles_m, P.tiles_n, kGroupN256, tm, tn);
        f32x4 acc[8][8];
        f256_accumulate<true>(P, tm * kBM, tn * kBN, pc.kbeg / kBK, pc.kend / kBK, smem, acc);
        if (pc.n > 1) {
            const int t = threadIdx.x;
            const __amdgpu_buffer_rsrc_t part =
                make_rsrc(sk.partials + (size_t)pc.region * pc.stride * (kF256SplitBytes / 4),
                          (long long)pc.stride * kF256SplitBytes);
            const int voff = 16 * t;
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    store_piece(part, voff, pc.q * kF256SplitBytes + 4096 * (8 * i + j),
                                acc[i][j]);
            __syncthreads();  // every wave is past its operand reads: LDS is free
            if (!splitk_arrive(sk, pc.region, reinterpret_cast<int*>(smem), pc.n)) return;
            // This workgroup's own piece is re-read like the others (the same
            // bits as its registers), so the sum runs in k order.  One row of
            // eight blocks at a time (the scheduling barriers keep the loads
            // of a row from being hoisted above the previous row's adds, which
            // would need registers for all 64 at once).
#pragma unroll
            for (int i = 0; i < 8; ++i) {
#pragma unroll
                for (int j = 0; j < 8; ++j) acc[i][j] = load_piece(part, voff, 4096 * (8 * i + j));
                __builtin_amdgcn_sched_barrier(0);
            }
            for (int q = 1; q < pc.n; ++q) {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
#pragma unroll
                    for (int j = 0; j < 8; ++j)
                        acc[i][j] += load_piece(part, voff, q * kF256SplitBytes + 4096 * (8 * i + j));
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        }
        f256_epilogue(P, tm * kBM, tn * kBN, smem, acc);
    });
}

// Workgroups of the stream-K kernel resident at once (one per CU).
long long f256_resident_workgroups() {
    static const long long n = [] {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sgemm_f32_256_sk, 256, 0) !=
                hipSuccess ||
            per_cu < 1) {
            (void)hipGetLastError();
            per_cu = 1;
        }
        return (long long)per_cu * device_cu_count();
    }();
    return n;
}

}  // namespace

bool f32_256_supported(const Problem& p) {
    // Byte offsets from a tile's first row of A and first column of B stay
    // below 2^31 for the whole k loop.
    const long long a_bytes = 4LL * ((long long)p.k * p.lda + kBM);
    const long long b_bytes = 4LL * ((long long)kBN * p.ldb + p.k);
    return columns_16b_aligned(p) && p.k % kBK == 0 && p.k >= 2 * kBK &&
           a_bytes < (1LL << 31) && b_bytes < (1LL << 31);
}

hipError_t launch_f32_256(const Problem& p, hipStream_t s, bool streamk, const Arena* arena) {
    Problem q = p;
    q.tiles_m = div_up(p.m, kBM);
    q.tiles_n = div_up(p.n, kBN);
    q.tile_begin = 0;
    const long long tiles = (long long)q.tiles_m * q.tiles_n;
    if (tiles <= 0) return hipSuccess;
    if (tiles > 0x7fffffffLL) return hipErrorInvalidValue;
    SplitK sk{nullptr, nullptr, 1, p.k, 0};
    long long nwg = 0;
    // The no-allocation entry point (arena) runs the whole tiles: its array
    // is far smaller than the stream-K partials of this tile.
    if (streamk && arena == nullptr) {
        const hipError_t e = streamk_plan(s, tiles, f256_resident_workgroups(), p.k, kBK,
                                          kF256SplitBytes, kF256StreamKPieceIters, &sk, &nwg,
                                          nullptr);
        if (e != hipSuccess) return e;
    }
    (void)hipGetLastError();  // the status read below must belong to this launch
    if (nwg > 0) {
        sk.dp_tiles = (int)(tiles - tiles % f256_resident_workgroups());
        hipLaunchKernelGGL(k_sgemm_f32_256_sk, dim3((unsigned)(sk.dp_tiles + nwg)), dim3(256), 0,
                           s, q, sk);
    } else {
        hipLaunchKernelGGL(k_sgemm_f32_256, dim3((unsigned)tiles), dim3(256), 0, s, q);
    }
    return hipGetLastError();
}

}  // namespace sgemm_amd
