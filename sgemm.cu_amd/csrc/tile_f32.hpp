// tile_f32.hpp -- the plain FP32 workgroup tile: one (BM x BN) block of C
// computed on v_mfma_f32_16x16x4_f32.  Shared by the single-problem kernels
// (sgemm_f32.hip) and the persistent grouped kernel (grouped.hip), so a tile
// produces the same bits whichever kernel runs it.
//
// Replaces the reference's FFMA kernels sgemm_128x256x8 / sgemm_128x128x8
// (src/kernels/128x256x8.cuh:4-387, src/kernels/128x128x8.cuh:4-386) with a
// wave64 design:
//
//  * The inner product runs on v_mfma_f32_16x16x4_f32: FP32 inputs, exact FP32
//    products, FP32 accumulation; each instruction is a rank-4 update of a
//    16x16 block held in 4 accumulator registers per lane.  A wave owns a
//    (BM/WM) x (BN/WN) register-blocked tile of such blocks.
//  * Global -> LDS: coalesced 16-byte loads into registers for k-tile t+1 while
//    the matrix cores work on tile t out of the other LDS stage (double
//    buffer, one barrier per k-tile).  A is kept m-contiguous in LDS (as in
//    HBM), B k-contiguous (as in HBM); row paddings make the fragment reads
//    bank-conflict free.
//  * Every edge is guarded: rows/cols/k outside the problem load as zero and
//    are never stored; ragged K is the last (zero-padded) k-tile.  Unlike the
//    reference (which uses n for ldb/ldc, 128x128x8.cuh:263,368), every leading
//    dimension is honoured.
//  * Fused epilogue: C = alpha*acc + beta*C; beta == 0 never reads C
//    (128x128x8.cuh:364-380 semantics).
//  * Barriers are executed by every wave of the workgroup (no barrier inside a
//    data-dependent branch, unlike 128x128x8.cuh:344-358).
#pragma once

#include <type_traits>

#include "kernels.hpp"

namespace sgemm_amd {

// AT_: keep A k-contiguous in LDS (As[m][k], like B) instead of m-contiguous
// (As[k][m], as in HBM).  Each thread then loads a 4(m) x 4(k) sub-block of A
// and transposes it in registers; in exchange every A fragment is one
// ds_read_b128 (four MFMA steps) instead of four 4-byte reads.
// WPE_: waves per SIMD the register allocation must allow.
// SCHED_: scheduling-hint pattern for full k-tiles (0: none; f32_schedule).
// PF_: k-tiles whose global loads are in flight at once.  1: the loads of
// k-tile t+1 are issued while t is multiplied and stored to LDS behind its
// first k-step.  PF > 1 keeps PF register stages, so the loads of t+PF are
// issued PF-1 k-tiles before they are stored: for grids with one workgroup
// per CU, where no other workgroup covers the memory latency.
// KG_: k groups.  KG x WM x WN waves; group g multiplies the 16-deep k-steps
// g, g+KG, ... of every k-tile into its own accumulators, and the groups'
// sums are added in group order through LDS before the epilogue.  Two groups
// put two waves of the tile on each SIMD, so one covers the other's LDS and
// barrier waits.  Whole tiles only (no split-K).
// CPF_: with beta != 0, a whole tile's C values are loaded into registers
// right after its first k-tile's operands, so the read of C overlaps the
// k loop instead of following it -- for short k, where C is most of the
// traffic and the loop is too short to hide a read issued in the epilogue.
// DL_: full k-tiles go from HBM straight into LDS (buffer_load_dwordx4 with
// the lds bit: each lane's 16 bytes land at the instruction's LDS base +
// 16 * lane), so no registers stage them and no LDS stores follow the loads;
// the wave waits for its own loads before the k-tile's closing barrier.  The
// placement is lane-contiguous, so the layout is chosen through which HBM
// address each lane reads.  A: blocks of 256 / BM k rows (one instruction,
// m-contiguous) with 1024 / BM floats of padding after each block (8 for BM
// 128, 16 for 64), so the four k rows 4 apart that one MFMA step reads start
// 16 banks apart.  B: rows
// of BK k-values per column with the 16-byte chunks of row n rotated by XOR
// with (n / 4) % 4 (BK 16) or (n / 2) % 8 (BK 32), so the 16 columns of a
// fragment read fall on distinct banks.  A k-tile that reaches past K goes
// through registers (masked) into the same layout, so the arithmetic and the
// result bits are those of the register-staged tile.  Needs 16-byte aligned
// operand columns (A, B 16-byte aligned, lda and ldb multiples of 4); the
// launcher picks the register-staged tile otherwise.
template <int BM_, int BN_, int BK_, int WM_, int WN_, bool AT_ = false, int WPE_ = 1,
          int SCHED_ = 0, int PF_ = 1, int KG_ = 1, bool CPF_ = false, bool DL_ = false>
struct F32Tile {
    static constexpr int BM = BM_, BN = BN_, BK = BK_, WM = WM_, WN = WN_;
    static constexpr bool AT = AT_;
    static constexpr int WPE = WPE_;
    static constexpr int SCHED = SCHED_;
    static constexpr int PF = PF_;
    static constexpr int KG = KG_;
    static constexpr bool CPF = CPF_;
    static constexpr bool DL = DL_;
    static constexpr int NT = 64 * WM * WN * KG;
    static constexpr int TM = BM / WM, TN = BN / WN;  // wave tile
    static constexpr int MB = TM / 16, NB = TN / 16;  // 16x16 MFMA blocks per wave
    // A layout: As[k][m] with row stride BM+4, As[m][k] with row stride BK+4
    // (AT), or blocks of ARB k rows with stride ABLK (DL).
    static constexpr int SA = AT ? BK + 4 : BM + 4;
    static constexpr int ARB = 256 / BM;               // DL: k rows per A block
    static constexpr int ABLK = 256 + 4 * ARB;         // DL: A block stride (floats)
    static constexpr int BRB = 256 / BK;               // DL: B columns per block
    static constexpr int A_FLOATS = DL ? (BK / ARB) * ABLK : AT ? BM * SA : BK * SA;
    static constexpr int SB = DL ? BK : BK + 4;       // Bs[n][k] row stride (floats)
    static constexpr int STAGE = A_FLOATS + BN * SB;  // floats per LDS stage
    static constexpr int LDS_FLOATS = 2 * STAGE;
    static constexpr int A_VEC = BM * BK / 4 / NT;  // float4 of A per thread per k-tile
    static constexpr int B_VEC = BN * BK / 4 / NT;
    static constexpr int A_KSTEP = NT / (BM / 4);   // k rows covered by one pass of threads
    static constexpr int B_NSTEP = NT / (BK / 4);   // n columns covered by one pass
    static_assert(TM % 16 == 0 && TN % 16 == 0, "wave tile must be 16-aligned");
    static_assert(BK % 16 == 0, "BK must be a multiple of 16");
    static_assert((BM * BK / 4) % NT == 0 && (BN * BK / 4) % NT == 0, "load split");
    static_assert(NT % (BM / 4) == 0 && NT % (BK / 4) == 0, "load mapping");
    static_assert(!AT || (A_VEC % 4 == 0 && (BM / 4) * (BK / 4) == NT * (A_VEC / 4)),
                  "transposed A needs whole 4x4 sub-blocks per thread");
    static_assert(LDS_FLOATS * 4 <= 160 * 1024, "LDS budget");
    static_assert(PF >= 1 && PF <= 4, "prefetch depth");
    static_assert(KG >= 1 && BK % (16 * KG) == 0, "k groups split the k-tile's k-steps");
    static_assert(KG == 1 || (KG - 1) * (TM / 16) * (TN / 16) * 64 * WM * WN * 4 <= LDS_FLOATS,
                  "k-group sums fit in LDS");
    static_assert(!CPF || KG == 1, "C prefetch on one-group tiles");
    static_assert(epilogue_lds_floats(WM * WN, TM, 16) <= LDS_FLOATS, "epilogue LDS");
    // DL: one load instruction fills one A block (ARB k rows of BM) or one B
    // block (BRB columns of BK); each wave issues A_VEC of the first kind and
    // B_VEC of the second per k-tile.
    static_assert(!DL || ((BM == 64 || BM == 128) && (BK == 16 || BK == 32) && !AT &&
                          PF == 1 && KG == 1 && !CPF && SCHED == 0 &&
                          A_VEC * WM * WN == BK / ARB && B_VEC * WM * WN == BN / BRB),
                  "direct-to-LDS tile shape");
};

// Instruction-group order for the compiler's scheduler in a full k-tile
// (__builtin_amdgcn_sched_group_barrier: mask, count, sync id; 0x008 MFMA,
// 0x002 VALU, 0x020 VMEM read, 0x100 DS read, 0x200 DS write).  Pattern 1:
// global loads first, all operand reads, then the MFMAs with one LDS store of
// the next k-tile after each of the later ones.  Pattern 2: as 1, but the
// operand reads are spread over the first MFMAs.
template <class T>
__device__ inline void f32_schedule() {
    constexpr int NM = (T::BK / 16) * 4 * T::MB * T::NB;            // MFMAs
    constexpr int GLOADS = T::A_VEC + T::B_VEC;                       // 16-byte loads
    constexpr int READS = (T::BK / 16) * ((T::AT ? 1 : 2) * T::MB + T::NB);
    constexpr int WRITES = T::A_VEC + T::B_VEC;                       // ds_write_b128
    constexpr int WSTART = NM / 4;
    __builtin_amdgcn_sched_group_barrier(0x020, GLOADS, 0);
    if constexpr (T::SCHED == 1) __builtin_amdgcn_sched_group_barrier(0x100, READS, 0);
    int w = 0, r = 0;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        if (T::SCHED == 2 && r < READS && m % 2 == 0) {
            __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
            r += 2;
        }
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        if (m >= WSTART && w < WRITES) {
            __builtin_amdgcn_sched_group_barrier(0x002, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
            ++w;
        }
    }
}

// PEEL: run the k-tiles whose successor is full through the branch-free
// body (faster, more registers); the persistent grouped kernel, which holds
// several tile shapes at once, turns it off to keep its register budget.
// Both orders do the same arithmetic, so results are identical bits.
template <class T, bool PEEL = true>
struct F32TileRunner {
    static constexpr int BM = T::BM, BN = T::BN, BK = T::BK;

    // Computes tile (tm, tn) of problem P.  `smem` holds T::LDS_FLOATS floats.
    // All threads of the workgroup must call this together.  With `sk`, only
    // split `split` of the k range runs here (see SplitK), and `slot` (unique
    // per split tile of the launch, below the workspace's tile count) picks
    // the tile's partials and arrival counter.
    __device__ static void run(const Problem& P, int tm, int tn, float* smem,
                               const SplitK* sk = nullptr, int split = 0, int slot = 0) {
        if (!sk) {
            run_piece(P, tm, tn, smem, nullptr, KPiece{0, P.k, 0, 1, 0, 1});
            return;
        }
        const int kbeg = split * sk->kchunk;
        run_piece(P, tm, tn, smem, sk,
                  KPiece{kbeg, min(P.k, kbeg + sk->kchunk), split, sk->splits, slot, sk->splits});
    }

    // As run, for the k range and reduction slot of `pc` (pc.n == 1: the
    // whole tile's k range, stored straight to C).
    __device__ static void run_piece(const Problem& P, int tm, int tn, float* smem,
                                     const SplitK* sk, const KPiece pc) {
        const int t = threadIdx.x;
        const int lane = t & 63, wave = t >> 6;
        constexpr int GW = T::WM * T::WN;  // waves per k group
        const int kg = T::KG == 1 ? 0 : wave / GW;
        const int wv = T::KG == 1 ? wave : wave % GW;
        const int wm = wv % T::WM, wn = wv / T::WM;
        const int l16 = lane & 15, g = lane >> 4;
        const int m0 = tm * BM, n0 = tn * BN;
        const int M = P.m, N = P.n, K = P.k;

        // ---- per-thread global load mapping (fixed across k-tiles).  Loads
        // go through buffer resources rebased per k-tile: rows >= M or
        // columns >= N read either 0 (past the operand's extent) or
        // in-extent values that only reach C rows/columns that are never
        // stored; k >= K is zeroed explicitly on both operands.
        // A, m-contiguous layout: float4 along m at (a_m, a_k + r*A_KSTEP).
        const int a_m = 4 * (t % (BM / 4));  // row inside the tile
        const int a_k = t / (BM / 4);        // first k row inside the tile
        // A, k-contiguous layout: 4x4 sub-block at rows 4*at_m4.., k 4*at_k4..;
        // a 16-lane group stores two rows 8 apart (disjoint banks at BK+4).
        constexpr int AS = BK / 4;
        const int at_k4 = t % AS;
        const int at_q = t / 16;
        const int at_m4 = (at_q % 2) + 2 * ((t / AS) % 2) + 4 * (at_q / 2);
        // B: float4 along k; BK/4 lanes cover one column.  Each 16-lane
        // group (one 256-byte LDS pass) writes rows spaced BK/4 apart, which
        // the BK+4 row stride puts on disjoint banks.
        constexpr int S = BK / 4, R = 16 / S;
        static_assert(16 % S == 0, "B store mapping");
        const int b_k = 4 * (t % S);
        const int b_q = t / 16;
        const int b_n = (b_q % S) + S * ((t / S) % R) + 16 * (b_q / S);
        // DL: load instruction r of this wave fills LDS block wave * A_VEC + r
        // of A (wave * B_VEC + r of B).  A: lane / (BM / 4) is the k row
        // within the block, lane % (BM / 4) the four m it reads.  B: lane /
        // (BK / 4) is the column within the block, lane % (BK / 4) the chunk
        // position, which holds k-chunk (lane % (BK / 4)) ^ dl_swz(column).
        const int dl_wave = __builtin_amdgcn_readfirstlane(wave);
        const int dl_arow = lane / (BM / 4), dl_am = 4 * (lane % (BM / 4));
        const int dl_col = lane / (BK / 4);
        auto dl_swz = [](int n) { return BK == 16 ? (n >> 2) & 3 : (n >> 1) & 7; };
        int a_off[T::A_VEC], b_off[T::B_VEC];
#pragma unroll
        for (int r = 0; r < T::A_VEC; ++r) {
            if constexpr (T::DL)
                a_off[r] = 4 * (dl_am + (T::ARB * (dl_wave * T::A_VEC + r) + dl_arow) * P.lda);
            else if constexpr (T::AT)
                a_off[r] = 4 * (4 * (at_m4 + (r / 4) * (T::NT / AS)) +
                                (4 * at_k4 + r % 4) * P.lda);
            else
                a_off[r] = 4 * (a_m + (a_k + r * T::A_KSTEP) * P.lda);
        }
#pragma unroll
        for (int r = 0; r < T::B_VEC; ++r) {
            if constexpr (T::DL) {
                const int n = T::BRB * (dl_wave * T::B_VEC + r) + dl_col;
                b_off[r] = 4 * (4 * ((lane % (BK / 4)) ^ dl_swz(n)) + n * P.ldb);
            }
            else
                b_off[r] = 4 * (b_k + (b_n + r * T::B_NSTEP) * P.ldb);
        }
        const long long a_ext = extent(M, K, P.lda), b_ext = extent(K, N, P.ldb);

        // Raw loads land in ra/rb (register stage k-tile % PF); the k >= K
        // masking happens in sstore, after the MFMAs of the current tile, so
        // nothing waits on the loads early.  Stage indices are compile-time
        // constants (std::integral_constant) so the stages stay in registers.
        f32x4 ra[T::PF][T::A_VEC], rb[T::PF][T::B_VEC];

        auto gload = [&](int k0, auto stage_tag) {
            constexpr int q = decltype(stage_tag)::value;
            const long long a_base = m0 + (long long)k0 * P.lda;
            const long long b_base = k0 + (long long)n0 * P.ldb;
            const auto ar = make_rsrc(P.A + a_base, 4 * (a_ext - a_base));
            const auto br = make_rsrc(P.B + b_base, 4 * (b_ext - b_base));
#pragma unroll
            for (int r = 0; r < T::A_VEC; ++r) ra[q][r] = bload4(ar, a_off[r]);
#pragma unroll
            for (int r = 0; r < T::B_VEC; ++r) rb[q][r] = bload4(br, b_off[r]);
        };
        // DL: a full k-tile straight into LDS `stage` (same addresses as gload).
        auto dma = [&](float* stage, int k0) {
            const long long a_base = m0 + (long long)k0 * P.lda;
            const long long b_base = k0 + (long long)n0 * P.ldb;
            const auto ar = make_rsrc(P.A + a_base, 4 * (a_ext - a_base));
            const auto br = make_rsrc(P.B + b_base, 4 * (b_ext - b_base));
#pragma unroll
            for (int r = 0; r < T::A_VEC; ++r)
                buffer_load_lds16(ar, stage + (dl_wave * T::A_VEC + r) * T::ABLK, a_off[r]);
#pragma unroll
            for (int r = 0; r < T::B_VEC; ++r)
                buffer_load_lds16(br, stage + T::A_FLOATS + (dl_wave * T::B_VEC + r) * 256, b_off[r]);
        };
        // MASK: zero k >= K (only the last k-tile can reach past K).
        auto sstore_impl = [&](float* stage, int k0, auto mask, auto stage_tag) {
            constexpr bool MASK = decltype(mask)::value;
            constexpr int q = decltype(stage_tag)::value;
            float* As = stage;
            float* Bs = stage + T::A_FLOATS;
            if constexpr (T::DL) {
                // The slots dma() would fill, k >= K zeroed.
#pragma unroll
                for (int r = 0; r < T::A_VEC; ++r) {
                    const bool ok =
                        !MASK || k0 + T::ARB * (dl_wave * T::A_VEC + r) + dl_arow < K;
                    *reinterpret_cast<f32x4*>(&As[(dl_wave * T::A_VEC + r) * T::ABLK + 4 * lane]) =
                        ok ? ra[q][r] : f32x4{0.f, 0.f, 0.f, 0.f};
                }
#pragma unroll
                for (int r = 0; r < T::B_VEC; ++r) {
                    const int n = T::BRB * (dl_wave * T::B_VEC + r) + dl_col;
                    const int kv = K - (k0 + 4 * ((lane % (BK / 4)) ^ dl_swz(n)));
                    f32x4 v;
#pragma unroll
                    for (int e = 0; e < 4; ++e) v[e] = !MASK || e < kv ? rb[q][r][e] : 0.f;
                    *reinterpret_cast<f32x4*>(&Bs[(dl_wave * T::B_VEC + r) * 256 + 4 * lane]) = v;
                }
                return;
            }
            if constexpr (T::AT) {
                const int ka = K - (k0 + 4 * at_k4);  // valid k in this sub-block
#pragma unroll
                for (int sb = 0; sb < T::A_VEC / 4; ++sb)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        f32x4 v;  // row 4*m4 + i, k = 4*at_k4 .. +3
#pragma unroll
                        for (int kk = 0; kk < 4; ++kk) v[kk] = !MASK || kk < ka ? ra[q][4 * sb + kk][i] : 0.f;
                        const int row = 4 * (at_m4 + sb * (T::NT / AS)) + i;
                        *reinterpret_cast<f32x4*>(&As[row * T::SA + 4 * at_k4]) = v;
                    }
            } else {
#pragma unroll
                for (int r = 0; r < T::A_VEC; ++r) {
                    const f32x4 v =
                        !MASK || (k0 + a_k + r * T::A_KSTEP < K) ? ra[q][r] : f32x4{0.f, 0.f, 0.f, 0.f};
                    *reinterpret_cast<f32x4*>(&As[(a_k + r * T::A_KSTEP) * T::SA + a_m]) = v;
                }
            }
            const int kv = K - (k0 + b_k);  // valid k in each B float4
#pragma unroll
            for (int r = 0; r < T::B_VEC; ++r) {
                f32x4 v;
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = !MASK || e < kv ? rb[q][r][e] : 0.f;
                *reinterpret_cast<f32x4*>(&Bs[(b_n + r * T::B_NSTEP) * T::SB + b_k]) = v;
            }
        };
        auto sstore = [&](float* stage, int k0, auto stage_tag) {
            if (k0 + BK <= K) sstore_impl(stage, k0, std::false_type{}, stage_tag);
            else sstore_impl(stage, k0, std::true_type{}, stage_tag);
        };
        using Stage0 = std::integral_constant<int, 0>;

        f32x4 acc[T::MB][T::NB];
#pragma unroll
        for (int i = 0; i < T::MB; ++i)
#pragma unroll
            for (int j = 0; j < T::NB; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

        const int kbeg = pc.kbeg;
        const int kend = pc.kend;
        const int ktiles = kend > kbeg ? div_up(kend - kbeg, BK) : 0;
        // C prefetch (whole tiles with beta != 0): issued after the first
        // k-tile's operand loads, in the epilogue's read-back mapping
        // (epilogue_block: lane -> column c_in of each pass, rows r..r+3).
        constexpr int LPC = T::TM / 4, CPP = 64 / LPC, PASSES = 16 / CPP;
        constexpr int NCP = T::CPF ? T::NB * PASSES : 1;
        f32x4 cpre[NCP];
        const bool prefetch_c = T::CPF && pc.n == 1 && P.beta != 0.f;
        auto load_c = [&]() {
            const int row = m0 + wm * T::TM + 4 * (lane % LPC);
            const int rvalid = min(max(M - row, 0), 4);
#pragma unroll
            for (int j = 0; j < T::NB; ++j)
#pragma unroll
                for (int p = 0; p < PASSES; ++p) {
                    const int col = n0 + wn * T::TN + 16 * j + p * CPP + lane / LPC;
                    cpre[(j * PASSES + p) % NCP] = load4(
                        P.C + (size_t)min(col, N - 1) * P.ldc + row, col < N ? rvalid : 0);
                }
        };
        if (T::DL && ktiles > 0 && kbeg + BK <= K) {
            dma(smem, kbeg);
            wait_vmem();
            __syncthreads();
        } else if (ktiles > 0) {
            gload(kbeg, Stage0{});
            if constexpr (T::CPF)
                if (prefetch_c) load_c();
            if constexpr (T::PF > 1) {
#pragma unroll
                for (int q = 1; q < T::PF; ++q) {
                    if (q < ktiles) {
                        // q is a loop constant after unrolling; dispatch it
                        // to the matching compile-time stage.
                        if (q == 1) gload(kbeg + BK, std::integral_constant<int, 1>{});
                        if constexpr (T::PF > 2)
                            if (q == 2) gload(kbeg + 2 * BK, std::integral_constant<int, 2 % T::PF>{});
                        if constexpr (T::PF > 3)
                            if (q == 3) gload(kbeg + 3 * BK, std::integral_constant<int, 3 % T::PF>{});
                    }
                }
            }
            sstore(smem, kbeg, Stage0{});
            __syncthreads();
        }
        // One k-tile: MFMAs out of the current stage while the next k-tile
        // is stored to the other one and k-tile kt+PF is loaded into the
        // register stage k-tile kt has just vacated.  NEXT_FULL: a next
        // k-tile exists and lies entirely below K, so the body has no
        // masking (and, at PF 1, no branch).  P = kt % PF.
        auto iteration = [&](int kt, auto next_full_tag, auto p_tag) {
            constexpr bool NEXT_FULL = decltype(next_full_tag)::value;
            constexpr int P = decltype(p_tag)::value;
            using NextStage = std::integral_constant<int, (P + 1) % T::PF>;
            const float* As = smem + (kt & 1) * T::STAGE;
            const float* Bs = As + T::A_FLOATS;
            float* next_stage = smem + ((kt + 1) & 1) * T::STAGE;
            const bool more = NEXT_FULL || kt + 1 < ktiles;
            const int knext = kbeg + (kt + 1) * BK;
            if constexpr (T::DL && NEXT_FULL) dma(next_stage, knext);
            else if (T::PF == 1 ? more : kt + T::PF < ktiles) gload(kbeg + (kt + T::PF) * BK, p_tag);
            if constexpr (NEXT_FULL && T::SCHED == 0) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int kq = 0; kq < BK / 16 / T::KG; ++kq) {
                const int kk = T::KG == 1 ? 16 * kq : 16 * (kq * T::KG + kg);
                // Lane group g supplies k = kk + 4g + s at MFMA step s, for A
                // and B alike, so B's four steps are one 16-byte LDS read.
                f32x4 a[T::MB];
                f32x4 b[T::NB];
#pragma unroll
                for (int i = 0; i < T::MB; ++i) {
                    if constexpr (T::DL) {
#pragma unroll
                        for (int s = 0; s < 4; ++s)
                            a[i][s] = As[((kk + 4 * g + s) / T::ARB) * T::ABLK +
                                         ((kk + 4 * g + s) % T::ARB) * BM + wm * T::TM +
                                         16 * i + l16];
                    } else if constexpr (T::AT) {
                        a[i] = *reinterpret_cast<const f32x4*>(
                            &As[(wm * T::TM + 16 * i + l16) * T::SA + kk + 4 * g]);
                    } else {
#pragma unroll
                        for (int s = 0; s < 4; ++s)
                            a[i][s] = As[(kk + 4 * g + s) * T::SA + wm * T::TM + 16 * i + l16];
                    }
                }
#pragma unroll
                for (int j = 0; j < T::NB; ++j) {
                    if constexpr (T::DL) {
                        const int n = wn * T::TN + 16 * j + l16;
                        b[j] = *reinterpret_cast<const f32x4*>(
                            &Bs[n * T::SB + 4 * ((kk / 4 + g) ^ dl_swz(n))]);
                    }
                    else
                        b[j] = *reinterpret_cast<const f32x4*>(
                            &Bs[(wn * T::TN + 16 * j + l16) * T::SB + kk + 4 * g]);
                }
#pragma unroll
                for (int s = 0; s < 4; ++s)
#pragma unroll
                    for (int i = 0; i < T::MB; ++i)
#pragma unroll
                        for (int j = 0; j < T::NB; ++j)
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][s], b[j][s],
                                                                             acc[i][j], 0, 0, 0);
                // Stage the next k-tile behind the first k-step's MFMAs, so
                // the wait on its loads, the masking and the LDS stores
                // overlap matrix-core work instead of following all of it.
                if (kq == 0) {
                    if constexpr (NEXT_FULL) {
                        if constexpr (!T::DL) {
                            if constexpr (T::SCHED == 0) __builtin_amdgcn_sched_barrier(0);
                            sstore_impl(next_stage, knext, std::false_type{}, NextStage{});
                            if constexpr (T::SCHED == 0) __builtin_amdgcn_sched_barrier(0);
                        }
                    } else if (more) {
                        sstore(next_stage, knext, NextStage{});
                    }
                }
            }
            if constexpr (NEXT_FULL && T::SCHED != 0) f32_schedule<T>();
            // DL: this wave's loads of the next k-tile have reached LDS
            // before the caller's barrier.
            if constexpr (T::DL && NEXT_FULL) wait_vmem();
        };
        // k-tiles whose successor is full run the branch-free body, the last
        // one or two the general body (two loops in sequence, so the
        // accumulators never merge from two paths).
        const int kfull = K - kbeg >= BK ? (K - kbeg) / BK : 0;
        const int nfast = min(ktiles, kfull) - 1;
        int kt = 0;
        if constexpr (PEEL) {
            // PF k-tiles per pass, so each one's register stage is a constant.
            for (; kt + T::PF <= nfast; kt += T::PF) {
                iteration(kt, std::true_type{}, Stage0{});
                __syncthreads();
                if constexpr (T::PF > 1) {
                    iteration(kt + 1, std::true_type{}, std::integral_constant<int, 1>{});
                    __syncthreads();
                }
                if constexpr (T::PF > 2) {
                    iteration(kt + 2, std::true_type{}, std::integral_constant<int, 2 % T::PF>{});
                    __syncthreads();
                }
                if constexpr (T::PF > 3) {
                    iteration(kt + 3, std::true_type{}, std::integral_constant<int, 3 % T::PF>{});
                    __syncthreads();
                }
            }
        }
        for (; kt < ktiles; ++kt) {
            const int p = kt % T::PF;
            if (p == 0) iteration(kt, std::false_type{}, Stage0{});
            if constexpr (T::PF > 1)
                if (p == 1) iteration(kt, std::false_type{}, std::integral_constant<int, 1>{});
            if constexpr (T::PF > 2)
                if (p == 2) iteration(kt, std::false_type{}, std::integral_constant<int, 2 % T::PF>{});
            if constexpr (T::PF > 3)
                if (p == 3) iteration(kt, std::false_type{}, std::integral_constant<int, 3 % T::PF>{});
            __syncthreads();
        }

        if constexpr (T::KG > 1) {
            // Groups 1.. hand their sums to group 0 through LDS (free after
            // the k loop's last barrier); group 0 adds them in group order.
            constexpr int GT = 64 * GW;
            f32x4* red = reinterpret_cast<f32x4*>(smem);
            const int tg = t % GT;
            if (kg > 0) {
#pragma unroll
                for (int i = 0; i < T::MB; ++i)
#pragma unroll
                    for (int j = 0; j < T::NB; ++j)
                        red[(((kg - 1) * T::MB + i) * T::NB + j) * GT + tg] = acc[i][j];
            }
            __syncthreads();
            if (kg == 0) {
                for (int q = 1; q < T::KG; ++q)
#pragma unroll
                    for (int i = 0; i < T::MB; ++i)
#pragma unroll
                        for (int j = 0; j < T::NB; ++j)
                            acc[i][j] += red[(((q - 1) * T::MB + i) * T::NB + j) * GT + tg];
            }
            __syncthreads();     // LDS is free again for a caller's next tile
            if (kg > 0) return;  // the epilogue below has no barrier
        }

        if (T::KG == 1 && sk && pc.n > 1) {
            // Dump the accumulators in lane order (coalesced 16-byte stores),
            // then let the last workgroup of this tile reduce its pieces in k
            // order.
            const int tile = pc.region;
            constexpr int NACC = T::MB * T::NB;
            constexpr int SPLIT_BYTES = NACC * T::NT * 16;
            const __amdgpu_buffer_rsrc_t part =
                make_rsrc(sk->partials + (size_t)tile * pc.stride * NACC * T::NT * 4,
                          (long long)pc.stride * SPLIT_BYTES);
#pragma unroll
            for (int i = 0; i < T::MB; ++i)
#pragma unroll
                for (int j = 0; j < T::NB; ++j)
                    store_agent(part, pc.q * SPLIT_BYTES + ((i * T::NB + j) * T::NT + t) * 16,
                                acc[i][j]);
            if (!splitk_arrive(*sk, tile, reinterpret_cast<int*>(smem), pc.n)) return;
            // Sum the partials in split order, this workgroup's own re-read
            // like the others (bit-identical to its registers), so the sum
            // reuses the accumulators' registers.  One split's loads are all
            // issued before any is waited on.
#pragma unroll
            for (int i = 0; i < T::MB; ++i)
#pragma unroll
                for (int j = 0; j < T::NB; ++j)
                    acc[i][j] = load_agent(part, ((i * T::NB + j) * T::NT + t) * 16);
            for (int q = 1; q < pc.n; ++q) {
#pragma unroll
                for (int i = 0; i < T::MB; ++i)
#pragma unroll
                    for (int j = 0; j < T::NB; ++j)
                        acc[i][j] += load_agent(
                            part, q * SPLIT_BYTES + ((i * T::NB + j) * T::NT + t) * 16);
            }
        }

        // ---- epilogue through LDS (epilogue_block): whole 16-byte pieces
        // down each column of C.  The k loop's last barrier (or the split
        // reduction's) has freed the LDS.
        float* elds = smem + kEpilogueLdsSkip + wv * 16 * (T::TM + 4);
#pragma unroll
        for (int j = 0; j < T::NB; ++j)
            epilogue_block<T::TM, 16, T::MB>(
                elds, P, m0 + wm * T::TM, n0 + wn * T::TN + 16 * j,
                [&](int i, int) { return acc[i][j]; },
                [&](int p, const float* cc, int valid) -> f32x4 {
                    if constexpr (T::CPF)
                        if (prefetch_c && ktiles > 0) return cpre[(j * PASSES + p) % NCP];
                    return load4(cc, valid);
                });
    }
};

// ---------------------------------------------------------------- configs
// Config ids of the plain FP32 form (the order is the ABI's config id).
using F32_128x128 = F32Tile<128, 128, 32, 2, 2>;
// 64x64 runs one workgroup per CU on grids of up to 256 tiles, so its loads
// are issued two k-tiles ahead: 2048^3 118 -> 125, 3072^3 122 -> 131 TFLOP/s.
using F32_64x64 = F32Tile<64, 64, 32, 2, 2, false, 1, 0, 2>;
using F32_128x64 = F32Tile<128, 64, 32, 2, 2>;
using F32_256x128 = F32Tile<256, 128, 16, 2, 2>;
using F32_128x128_at = F32Tile<128, 128, 32, 2, 2, true, 2>;
using F32_128x128_k16 = F32Tile<128, 128, 16, 2, 2, false, 3, 1>;
// Config 5's tile with direct-to-LDS operand loads: no staging registers,
// 33 KB of LDS, four workgroups per CU (8192^3 140 -> 145, 16384^3 141.5 ->
// 147, 4096^3 134 -> 141 TFLOP/s, same result bits).  The launcher runs it
// for configs 5, 9 and 10 when the operand columns are 16-byte aligned.
using F32_128x128_k16_dl = F32Tile<128, 128, 16, 2, 2, false, 4, 0, 1, 1, false, true>;
// Config 1/7's 64x64x32 tile with direct-to-LDS loads (34 KB of LDS, four
// workgroups per CU; same result bits as the register-staged tile, which
// loads two k-tiles ahead): 2048^3 121 -> 129, 3072^3 127.5 -> 133 TFLOP/s.
using F32_64x64_dl = F32Tile<64, 64, 32, 2, 2, false, 4, 0, 1, 1, false, true>;
// 64x64x64 with two k groups of four waves, loads two k-tiles ahead: for
// grids of one 64x64 tile per CU (1024^3: 86 -> 92 TFLOP/s).
using F32_64x64_kg2 = F32Tile<64, 64, 64, 2, 2, false, 1, 0, 2, 2>;
// Short k (a few k-tiles): C prefetched at the start of the tile, so its
// read overlaps the k loop.  (Loading two k-tiles ahead as well spills on
// these tiles: 128x64 27, 128x128 84 VGPRs even without the C registers.)
// The 128x64 tile takes 16-deep k-tiles: half the LDS (27 KB), four
// workgroups per CU instead of three, so more tiles' C traffic is in flight
// at once (8192x8192x127 95 -> 98, x64 67.5 -> 72 TFLOP/s).
using F32_128x128_short = F32Tile<128, 128, 32, 2, 2, false, 2, 0, 1, 1, true>;
using F32_128x64_short = F32Tile<128, 64, 16, 2, 2, false, 4, 0, 1, 1, true>;

template <class T>
struct TypeTag {
    using type = T;
};

// Calls f(TypeTag<Tile>{}) for config `id`; returns false for an unknown id.
template <class F>
__host__ __device__ inline bool dispatch_f32(int id, F&& f) {
    switch (id) {
        case 0: f(TypeTag<F32_128x128>{}); return true;
        case 1: f(TypeTag<F32_64x64>{}); return true;
        case 2: f(TypeTag<F32_128x64>{}); return true;
        case 3: f(TypeTag<F32_256x128>{}); return true;
        case 4: f(TypeTag<F32_128x128_at>{}); return true;
        case 5: f(TypeTag<F32_128x128_k16>{}); return true;
        case 13: f(TypeTag<F32_128x128_short>{}); return true;
        case 14: f(TypeTag<F32_128x64_short>{}); return true;
        default: return false;
    }
}

}  // namespace sgemm_amd
