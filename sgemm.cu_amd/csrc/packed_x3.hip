// packed_x3.hip -- BF16x3 configs 12 and 13: the operands are split once, in a
// pass of their own, into hi/lo planes laid out the way the MFMAs read them,
// and a 256x256 tile multiplies those planes straight out of LDS (config 13
// adds config 11's stream-K tail).  No reference counterpart (the reference
// is FP32 FFMA only).
//
// Why.  The in-register split of the other configs (tile_x3.hpp) splits every
// element of A once per workgroup that loads it -- n / 256 times over the
// launch -- and that VALU work, with the LDS stores it feeds, is the largest
// cost of the 256x256 tile after the MFMAs.  Here each element is split once
// (about 5 TB/s of read + write for the pass), and the tile's k loop is
// direct-to-LDS loads, LDS reads and MFMAs only.
//
// Packed layout.  An operand is a set of rows (A: its rows; B: its columns)
// by k.  Rows are cut into panels of 256, k into k-tiles of 32; k-tile kt of
// panel p is 32 KB: 16 blocks of 16 rows, each a 1 KB hi chunk then a 1 KB lo
// chunk.  Chunk lane l (16 bytes) holds row l % 16 of the block at k = 8 * (l
// / 16) .. +7 -- exactly the operand of lane l of v_mfma_f32_16x16x32_bf16
// -- so every LDS read of an operand is one contiguous 1 KB ds_read_b128 and
// every load is one contiguous 1 KB buffer_load_dwordx4 ... lds.  Rows past
// the operand and k past K are zero.  Panel p's k-tiles follow each other.
//
// Results.  The split is split4 of tile_x3.hpp and every accumulator takes
// al*bh, ah*bl, ah*bh per k-tile in k order, as config 2 does, so config 12
// gives config 2's bits and config 13 config 11's.  Where the packed planes
// cannot be had -- a stream under capture whose pack workspace would have to
// grow, the no-allocation entry point, a plane past 32-bit offsets or the
// size cap, a failed allocation -- the launch runs config 2 (or 11) instead,
// with the same result.
#include <mutex>

#include "tile_f32.hpp"
#include "tile_x3.hpp"

namespace sgemm_amd {

namespace {

constexpr int kPanel = 256;                 // rows per panel
constexpr int kPk = 32;                     // k per k-tile
constexpr int kPanelTileBytes = 32 * 1024;  // one panel's k-tile: 16 x {hi, lo} x 1 KB
// Largest pack workspace (both planes) a stream may hold; larger problems run
// config 2.
constexpr size_t kPackCapBytes = size_t(16) << 30;

__host__ __device__ inline size_t chunk_offset(int panel, int ktiles, int kt, int blk) {
    return (((size_t)panel * ktiles + kt) * 16 + blk) * 2048;
}

// Bytes of one packed plane pair for `rows` rows and k.
size_t packed_bytes(int rows, int k) {
    return (size_t)div_up(rows, kPanel) * div_up(k, kPk) * kPanelTileBytes;
}

__device__ inline void store_chunk_lane(char* o, f32x4 x0, f32x4 x1) {
    bf16x4_t h0, l0, h1, l1;
    split4(x0, h0, l0);
    split4(x1, h1, l1);
    const bf16x8_t h = {h0[0], h0[1], h0[2], h0[3], h1[0], h1[1], h1[2], h1[3]};
    const bf16x8_t l = {l0[0], l0[1], l0[2], l0[3], l1[0], l1[1], l1[2], l1[3]};
    *reinterpret_cast<bf16x8_t*>(o) = h;
    *reinterpret_cast<bf16x8_t*>(o + 1024) = l;
}

// Operand whose rows are contiguous (A: element (r, k) at X[r + k * ld]).
// Workgroup (kt, panel), 256 threads.  The 256 x 32 block is read one column
// (1 KB) per wave instruction into LDS [k][row]; then wave w packs blocks w,
// w + 4, w + 8, w + 12, lane l taking row l % 16 of the block and k group
// l / 16, so each store writes one whole 1 KB chunk.
constexpr int kTransposeStride = kPanel + 4;  // floats per k row of the LDS block
__global__ __launch_bounds__(256) void k_x3_pack_rows(const float* __restrict__ X, int rows, int K,
                                                      int ld, char* __restrict__ out, int ktiles) {
    __shared__ __attribute__((aligned(16))) float blk_lds[kPk * kTransposeStride];
    const int kt = blockIdx.x, panel = blockIdx.y, t = threadIdx.x;
    const int r4 = 4 * (t % 64), kq = t / 64;
    const int k0 = kt * kPk;
    // One buffer resource per workgroup (a per-lane base would make the
    // compiler loop over the distinct bases); per-lane offsets below 2^31
    // for ld <= kMaxLd.
    const long long base = (long long)panel * kPanel + (long long)k0 * ld;
    const auto rs = make_rsrc(X + base, 4 * (extent(rows, K, ld) - base));
    const int rvalid = rows - (panel * kPanel + r4);
    f32x4 v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = bload4(rs, 4 * (r4 + (kq + 4 * e) * ld));
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int k = kq + 4 * e;
        f32x4 x;
#pragma unroll
        for (int i = 0; i < 4; ++i) x[i] = (i < rvalid && k0 + k < K) ? v[e][i] : 0.f;
        *reinterpret_cast<f32x4*>(&blk_lds[k * kTransposeStride + r4]) = x;
    }
    __syncthreads();
    const int lane = t & 63, w = t >> 6;
    const int lr = lane % 16, g = lane / 16;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int blk = w + 4 * b;
        f32x4 x0, x1;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            x0[e] = blk_lds[(8 * g + e) * kTransposeStride + blk * 16 + lr];
            x1[e] = blk_lds[(8 * g + 4 + e) * kTransposeStride + blk * 16 + lr];
        }
        store_chunk_lane(out + chunk_offset(panel, ktiles, kt, blk) + lane * 16, x0, x1);
    }
}

// Operand whose k is contiguous (B: element (k, c) at X[k + c * ld], c the
// packed row).  Thread t, pass s: row c = 64 s + t / 4 of the panel, k =
// 8 (t % 4) .. +7 of the k-tile (two 16-byte loads); the wave's 64 lanes fill
// one whole chunk per store.
__global__ __launch_bounds__(256) void k_x3_pack_cols(const float* __restrict__ X, int cols, int K,
                                                      int ld, char* __restrict__ out, int ktiles) {
    const int kt = blockIdx.x, panel = blockIdx.y, t = threadIdx.x;
    const int g = t % 4;
    const int k0 = kt * kPk + 8 * g;
    const long long base = (long long)kt * kPk + (long long)panel * kPanel * ld;
    const auto rs = make_rsrc(X + base, 4 * (extent(K, cols, ld) - base));
    f32x4 a[4], b[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int off = 4 * (8 * g + (s * 64 + t / 4) * ld);
        a[s] = bload4(rs, off);
        b[s] = bload4(rs, off + 16);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int cc = s * 64 + t / 4;
        const bool cvalid = panel * kPanel + cc < cols;
        f32x4 x0, x1;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            x0[e] = (cvalid && k0 + e < K) ? a[s][e] : 0.f;
            x1[e] = (cvalid && k0 + 4 + e < K) ? b[s][e] : 0.f;
        }
        const int lane = (cc % 16) + 16 * g;
        store_chunk_lane(out + chunk_offset(panel, ktiles, kt, cc / 16) + lane * 16, x0, x1);
    }
}

// The packed tile: 256 x 256, eight waves (2 x 4) of 128 x 64, i.e. 8 x 4
// MFMA blocks per wave with one accumulator set; two 64 KB LDS stages (A's
// 32 KB, then B's).  Waves 0-3 load the A chunks of a k-tile, 8 each, waves
// 4-7 the B chunks.  k-tile t: read B's operands and A's first row, run the
// rows of MFMAs with the next row's A operands read during each (regions
// fenced by sched_barrier so the reads stay ahead); after row 0 issue the
// loads of k-tile t + 1 into the other stage (that stage's readers finished
// at the barrier that ended k-tile t - 1; issuing them after the k-tile's
// first LDS reads rather than before ran 513 -> 527 TFLOP/s at 8192^3); then
// wait for them and barrier.
constexpr int kPkNt = 512, kPkTm = 128, kPkTn = 64, kPkMb = 8, kPkNb = 4;
constexpr int kPkStage = 2 * kPanelTileBytes;
constexpr int kPkLds = 2 * kPkStage;
static_assert(epilogue_lds_floats(8, kPkTm, 16) * 4 <= kPkLds, "epilogue LDS");
// A cut tile's partials: 32 float4 accumulators x 512 threads per piece, in
// the layout of config 2's split pieces (tile_x3.hpp), so a stream-K tile is
// summed to the same bits as config 11's.
constexpr int kPkAcc = kPkMb * kPkNb;
constexpr int kPkSplitBytes = kPkAcc * kPkNt * 16;
static_assert(kPkSplitBytes == (int)kX3TileSplitBytes, "config 2's split layout");

typedef f32x4 PkAcc[kPkMb][kPkNb];

// Accumulates k-tiles [kt0, kt1) of tile (tm, tn) into acc (zeroed here).
// All threads of the workgroup call it; on return the LDS is free (a barrier
// ends every k-tile).
__device__ inline void packed_accumulate(int tm, int tn, const char* __restrict__ pa,
                                         const char* __restrict__ pb, int ktiles, int kt0,
                                         int kt1, char* smem, PkAcc& acc) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave % 2, wn = wave / 2;
    const int dw = __builtin_amdgcn_readfirstlane(wave);
    const char* src = dw < 4 ? pa + (size_t)tm * ktiles * kPanelTileBytes
                             : pb + (size_t)tn * ktiles * kPanelTileBytes;
    const auto rs = make_rsrc(src, (long long)ktiles * kPanelTileBytes);
    const int cbase = 8 * (dw % 4);
    const int lbase = (dw < 4 ? 0 : kPanelTileBytes) + cbase * 1024;
    auto dma = [&](int stage, int kt) {
#pragma unroll
        for (int r = 0; r < 8; ++r)
            buffer_load_lds16(rs, reinterpret_cast<float*>(smem + stage * kPkStage + lbase + r * 1024),
                              kt * kPanelTileBytes + (cbase + r) * 1024 + lane * 16);
    };
    auto rd = [&](const char* p) { return *reinterpret_cast<const bf16x8_t*>(p + lane * 16); };
#pragma unroll
    for (int i = 0; i < kPkMb; ++i)
#pragma unroll
        for (int j = 0; j < kPkNb; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    if (kt0 < kt1) {
        dma(kt0 & 1, kt0);
        wait_vmem();
        __syncthreads();
    }
    for (int kt = kt0; kt < kt1; ++kt) {
        const char* As = smem + (kt & 1) * kPkStage;
        const char* Bs = As + kPanelTileBytes;
        const bool more = kt + 1 < kt1;
        bf16x8_t bh[kPkNb], bl[kPkNb], ah[2], al[2];
#pragma unroll
        for (int j = 0; j < kPkNb; ++j) {
            bh[j] = rd(Bs + ((wn * kPkNb + j) * 2) * 1024);
            bl[j] = rd(Bs + ((wn * kPkNb + j) * 2 + 1) * 1024);
        }
        ah[0] = rd(As + ((wm * kPkMb) * 2) * 1024);
        al[0] = rd(As + ((wm * kPkMb) * 2 + 1) * 1024);
#pragma unroll
        for (int i = 0; i < kPkMb; ++i) {
            const int c = i & 1, n = c ^ 1;
            if (i + 1 < kPkMb) {
                ah[n] = rd(As + ((wm * kPkMb + i + 1) * 2) * 1024);
                al[n] = rd(As + ((wm * kPkMb + i + 1) * 2 + 1) * 1024);
            }
#pragma unroll
            for (int j = 0; j < kPkNb; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[c], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
            for (int j = 0; j < kPkNb; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[c], bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
            for (int j = 0; j < kPkNb; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[c], bh[j], acc[i][j], 0, 0, 0);
            if (i == 0 && more) dma((kt + 1) & 1, kt + 1);
            __builtin_amdgcn_sched_barrier(0);
        }
        wait_vmem();
        __syncthreads();
    }
}

// Epilogue through LDS, as the other BF16x3 tiles (epilogue_block).  The LDS
// must be free; its first kEpilogueLdsSkip floats are left alone.
__device__ inline void packed_epilogue(const Problem& P, int tm, int tn, char* smem, PkAcc& acc) {
    const int wave = threadIdx.x >> 6;
    const int wm = wave % 2, wn = wave / 2;
    const int m0 = tm * 256, n0 = tn * 256;
    float* elds = reinterpret_cast<float*>(smem) + kEpilogueLdsSkip + wave * 16 * (kPkTm + 4);
#pragma unroll
    for (int j = 0; j < kPkNb; ++j)
        epilogue_block<kPkTm, 16, kPkMb>(
            elds, P, m0 + wm * kPkTm, n0 + wn * kPkTn + 16 * j,
            [&](int i, int) { return acc[i][j]; },
            [&](int, const float* cc, int valid) { return load4(cc, valid); });
}

__global__ __launch_bounds__(kPkNt) __attribute__((amdgpu_waves_per_eu(1)))
void k_sgemm_bf16x3_packed(Problem P, const char* __restrict__ pa, const char* __restrict__ pb,
                           int ktiles) {
    __shared__ __attribute__((aligned(16))) char smem[kPkLds];
    int tm, tn;
    tile_coords(xcd_remap(blockIdx.x, gridDim.x), P.tiles_m, P.tiles_n, 8, tm, tn);
    PkAcc acc;
    packed_accumulate(tm, tn, pa, pb, ktiles, 0, ktiles, smem, acc);
    packed_epilogue(P, tm, tn, smem, acc);
}

// Config 13: the packed tile with config 11's stream-K tail.  Workgroups
// [0, sk.dp_tiles) compute whole tiles; the rest run the k-tile ranges of
// streamk_pieces over the remaining tiles.  A cut tile's pieces store their
// partials in config 2's split layout and the last to arrive adds them in k
// order, as tile_x3.hpp does, so the result is config 11's bits.
__global__ __launch_bounds__(kPkNt) __attribute__((amdgpu_waves_per_eu(1)))
void k_sgemm_bf16x3_packed_sk(Problem P, SplitK sk, const char* __restrict__ pa,
                              const char* __restrict__ pb, int ktiles) {
    __shared__ __attribute__((aligned(16))) char smem[kPkLds];
    int tm, tn;
    const int b = blockIdx.x;
    if (b < sk.dp_tiles) {
        tile_coords(xcd_remap(b, sk.dp_tiles), P.tiles_m, P.tiles_n, 8, tm, tn);
        PkAcc acc;
        packed_accumulate(tm, tn, pa, pb, ktiles, 0, ktiles, smem, acc);
        packed_epilogue(P, tm, tn, smem, acc);
        return;
    }
    streamk_pieces(P, sk, b, kPk, [&](int tile, const KPiece& pc) {
        tile_coords(tile, P.tiles_m, P.tiles_n, 8, tm, tn);
        PkAcc acc;
        packed_accumulate(tm, tn, pa, pb, ktiles, pc.kbeg / kPk, div_up(pc.kend, kPk), smem, acc);
        if (pc.n > 1) {
            const int t = threadIdx.x;
            const __amdgpu_buffer_rsrc_t part =
                make_rsrc(sk.partials + (size_t)pc.region * pc.stride * kPkAcc * kPkNt * 4,
                          (long long)pc.stride * kPkSplitBytes);
            auto off = [&](int i, int j) { return ((i * kPkNb + j) * kPkNt + t) * 16; };
#pragma unroll
            for (int i = 0; i < kPkMb; ++i)
#pragma unroll
                for (int j = 0; j < kPkNb; ++j)
                    store_agent(part, pc.q * kPkSplitBytes + off(i, j), acc[i][j]);
            if (!splitk_arrive(sk, pc.region, reinterpret_cast<int*>(smem), pc.n)) return;
            // This workgroup's own piece is re-read like the others (the same
            // bits as its registers), so the sum runs in k order.
#pragma unroll
            for (int i = 0; i < kPkMb; ++i)
#pragma unroll
                for (int j = 0; j < kPkNb; ++j) acc[i][j] = load_agent(part, off(i, j));
            for (int q = 1; q < pc.n; ++q) {
#pragma unroll
                for (int i = 0; i < kPkMb; ++i)
#pragma unroll
                    for (int j = 0; j < kPkNb; ++j)
                        acc[i][j] += load_agent(part, q * kPkSplitBytes + off(i, j));
            }
        }
        packed_epilogue(P, tm, tn, smem, acc);
    });
}

// Workgroups of the stream-K kernel resident at once (one per CU).
long long packed_resident_workgroups() {
    static const long long n = [] {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sgemm_bf16x3_packed_sk, kPkNt,
                                                         0) != hipSuccess ||
            per_cu < 1) {
            (void)hipGetLastError();
            per_cu = 1;
        }
        return (long long)per_cu * device_cu_count();
    }();
    return n;
}

// The packed planes of p in the stream's workspace, launched; false (nothing
// launched) where they cannot be had.
bool launch_packs(const Problem& p, hipStream_t s, const Arena* arena, const char** pa,
                  const char** pb) {
    const int ktiles = div_up(p.k, kPk);
    const size_t bytes_a = packed_bytes(p.m, p.k), bytes_b = packed_bytes(p.n, p.k);
    const long long tiles = (long long)div_up(p.m, 256) * div_up(p.n, 256);
    if (arena != nullptr || tiles <= 0 || tiles > 0x7fffffffLL ||
        (long long)ktiles * kPanelTileBytes > 0x7fffffffLL || bytes_a + bytes_b > kPackCapBytes)
        return false;
    void* ws = nullptr;
    if (pack_workspace(s, bytes_a + bytes_b, &ws) != hipSuccess) return false;
    char* a = static_cast<char*>(ws);
    char* b = a + bytes_a;
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_x3_pack_rows, dim3(ktiles, div_up(p.m, kPanel)), dim3(256), 0, s, p.A,
                       p.m, p.k, p.lda, a, ktiles);
    hipLaunchKernelGGL(k_x3_pack_cols, dim3(ktiles, div_up(p.n, kPanel)), dim3(256), 0, s, p.B,
                       p.n, p.k, p.ldb, b, ktiles);
    *pa = a;
    *pb = b;
    return true;
}

}  // namespace

hipError_t launch_x3_packed(const Problem& p, hipStream_t s, const Arena* arena, bool streamk) {
    Problem q = p;
    q.tiles_m = div_up(p.m, 256);
    q.tiles_n = div_up(p.n, 256);
    q.tile_begin = 0;
    const long long tiles = (long long)q.tiles_m * q.tiles_n;
    if (tiles <= 0) return hipSuccess;
    // The stream-K plan first: its workspace (partials, counters) may fail
    // to grow under capture, which streamk_plan turns into whole tiles.
    SplitK sk{nullptr, nullptr, 1, p.k, 0};
    long long nwg = 0;
    if (streamk && arena == nullptr) {
        const hipError_t e = streamk_plan(s, tiles, packed_resident_workgroups(), p.k, kPk,
                                          kPkSplitBytes, kX3StreamKPieceIters, &sk, &nwg);
        if (e != hipSuccess) return e;
    }
    const char *pa = nullptr, *pb = nullptr;
    // The packs and the tile go onto the stream together: two host threads
    // calling on one stream must not interleave one call's packs between
    // the other's packs and tile, since both use the stream's planes.
    static std::mutex enqueue_mu;
    std::lock_guard<std::mutex> lock(enqueue_mu);
    // No packed planes (see the file comment): config 2 or 11, the same bits.
    if (!launch_packs(p, s, arena, &pa, &pb)) return launch_bf16x3(streamk ? 11 : 2, p, s, arena);
    const int ktiles = div_up(p.k, kPk);
    if (nwg > 0) {
        sk.dp_tiles = (int)(tiles - tiles % packed_resident_workgroups());
        hipLaunchKernelGGL(k_sgemm_bf16x3_packed_sk, dim3((unsigned)(sk.dp_tiles + nwg)),
                           dim3(kPkNt), 0, s, q, sk, pa, pb, ktiles);
    } else {
        hipLaunchKernelGGL(k_sgemm_bf16x3_packed, dim3((unsigned)tiles), dim3(kPkNt), 0, s, q, pa,
                           pb, ktiles);
    }
    return hipGetLastError();
}

}  // namespace sgemm_amd
