// tile_x3.hpp -- the BF16x3 split-precision workgroup tile on the gfx950 BF16
// matrix cores.  Shared by the single-problem kernels (sgemm_bf16x3.hip) and
// the persistent grouped kernel (grouped.hip).
//
// C := alpha*A*B + beta*C, column-major, FP32 in and out.  No reference
// counterpart (the reference is FP32 FFMA only); same operand conventions and
// epilogue semantics as sgemm_f32.hip.
//
// Each FP32 input x is split as it is staged into LDS:
//     hi = bf16_rne(x),  lo = bf16_rne(x - hi)      (x - hi is exact in FP32)
// (hi from x clamped to the largest finite bf16, see split4) and A*B is
// accumulated in FP32 as  hi*hi + hi*lo + lo*hi  on the BF16 MFMAs (BF16
// products are exact in FP32).  DUAL configs keep the small cross terms in
// accumulators of their own, added to the main ones in the epilogue; the
// others accumulate all three products in one set.  Dropped: lo*lo and the
// residual x - hi - lo, at most ~3*2^-16 |a||b| per product.  That residual
// (~2^-17 |x| for a two-term split) sets the error: on wide-range inputs
// the mean componentwise error is about 3x the plain FP32 kernel's whatever
// the accumulation order (a separate cross-term set or a fourth lo*lo
// product changes it by under 20 %).
//
// Global -> LDS: 16-byte loads of FP32 into registers for k-tile t+1 during
// the MFMAs of tile t; the split happens in registers; hi and lo land in LDS
// k-contiguous per row ([row][k], 16-byte padded rows) so every MFMA operand
// (8 consecutive k of one row per lane) is one ds_read_b128.  A is loaded in
// 4x4 (m x k) sub-blocks so its transpose to k-contiguous rows needs only
// 8-byte LDS writes.
#pragma once

#include <type_traits>

#include "kernels.hpp"

namespace sgemm_amd {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));

// BM x BN workgroup tile, BK k-tile, WM x WN waves.  MF selects the MFMA:
// 16 -> v_mfma_f32_16x16x32_bf16 (16x16 blocks, 4 accumulators per lane),
// 32 -> v_mfma_f32_32x32x16_bf16 (32x32 blocks, 16 accumulators per lane).
// DUAL keeps hi*lo + lo*hi in separate accumulators (added in the epilogue);
// otherwise all three products go into one set, which halves the accumulator
// registers and so allows larger wave tiles.  MID stages the next k-tile
// right after the first k-step's MFMAs instead of after the last one; it pays
// when two independent workgroups share a CU and hurts a lone wave per SIMD,
// whose loads then have less time to land.  WPE_ > 0 caps the registers so
// that many waves per SIMD fit (two 4-wave workgroups per CU: their barriers
// fall at different times, so one's MFMAs fill the other's bubbles).  SCHED_
// 1 runs full k-tiles through interleaved_iteration (X3TileRunner): the next
// k-tile's split spread over the MFMAs, its loads a k-tile ahead.
template <int BM_, int BN_, int BK_, int WM_, int WN_, int MF_ = 16, bool DUAL_ = true,
          bool MID_ = false, int WPE_ = 0, int SCHED_ = 0>
struct X3Tile {
    static constexpr int BM = BM_, BN = BN_, BK = BK_, WM = WM_, WN = WN_;
    static constexpr int MF = MF_;
    static constexpr bool DUAL = DUAL_;
    static constexpr bool MID = MID_;
    static constexpr int SCHED = SCHED_;
    // Waves per SIMD the kernel's registers must allow: two independent
    // workgroups per CU for the MID configs or when WPE_ asks for it, else
    // whatever the tile needs.
    static constexpr int WPE = WPE_ > 0 ? WPE_ : (MID_ ? 2 : 1);
    static constexpr int NT = 64 * WM * WN;
    static constexpr int TM = BM / WM, TN = BN / WN;
    static constexpr int MB = TM / MF, NB = TN / MF;     // MFMA blocks per wave
    static constexpr int KSTEP = MF == 16 ? 32 : 16;     // k per MFMA
    static constexpr int RS = BK * 2 + 16;               // LDS row stride, bytes
    static constexpr int A_BYTES = BM * RS;              // one of Ahi / Alo
    static constexpr int B_BYTES = BN * RS;
    static constexpr int STAGE = 2 * A_BYTES + 2 * B_BYTES;
    static constexpr int LDS_BYTES = 2 * STAGE;
    static constexpr int A_SUB = (BM / 4) * (BK / 4) / NT;  // 4x4 sub-blocks of A per thread
    static constexpr int A_MSTEP = NT / (BK / 4);           // m4 rows per pass
    static constexpr int B_VEC = BN * BK / 4 / NT;          // float4 of B per thread
    static constexpr int B_NSTEP = NT / (BK / 4);
    static_assert(MF == 16 || MF == 32, "MFMA shape");
    static_assert(BK % KSTEP == 0 && BK % 16 == 0, "BK must be a multiple of the MFMA k");
    static_assert(TM % MF == 0 && TN % MF == 0, "wave tile");
    static_assert(NT % (BK / 4) == 0, "load mapping");
    static_assert(A_SUB >= 1 && (BM / 4) * (BK / 4) % NT == 0, "A split");
    static_assert(B_VEC >= 1 && (BN * BK / 4) % NT == 0, "B split");
    static_assert(LDS_BYTES <= 160 * 1024, "LDS budget");
    static_assert(epilogue_lds_floats(WM * WN, TM, MF) * 4 <= LDS_BYTES, "epilogue LDS");
};

// hi/lo split of four values (round to nearest even both times), two at a
// time so each rounding is one v_cvt_pk_bf16_f32 for a pair.
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

// Largest finite bf16 (0x7f7f), exactly representable in FP32.
constexpr float kBf16Max = 3.38953139e38f;

// hi is rounded from x clamped to +-kBf16Max (one v_med3_f32, the identity
// for every |x| <= kBf16Max), so it is always finite:
//  * a finite |x| above kBf16Max, which would round to Inf, keeps hi =
//    +-kBf16Max and lo = bf16(x - hi), and x - hi is still exact (Sterbenz);
//  * x = +-Inf gives lo = +-Inf, so lo*b_hi carries the infinity with the
//    sign of the FP32 product, and Inf * 0 turns into NaN only where b
//    itself is 0, as in FP32.  (With hi = Inf and lo = 0 instead, hi*b_lo
//    would be Inf * 0 = NaN for every bf16-exact b.);
//  * NaN gives lo = NaN.
__device__ inline void split4(f32x4 x, bf16x4_t& hi, bf16x4_t& lo) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const f32x2 v = {x[2 * p], x[2 * p + 1]};
        const f32x2 c = {__builtin_amdgcn_fmed3f(v[0], -kBf16Max, kBf16Max),
                         __builtin_amdgcn_fmed3f(v[1], -kBf16Max, kBf16Max)};
        const bf16x2_t h = __builtin_convertvector(c, bf16x2_t);
        const f32x2 r = v - __builtin_convertvector(h, f32x2);  // exact in FP32
        const bf16x2_t l = __builtin_convertvector(r, bf16x2_t);
        hi[2 * p] = h[0];
        hi[2 * p + 1] = h[1];
        lo[2 * p] = l[0];
        lo[2 * p + 1] = l[1];
    }
}

// PEEL: run the k-tiles whose successor is full through the branch-free
// body (faster, more registers); the persistent grouped kernel, which holds
// several tile shapes at once, turns it off to keep its register budget.
// Both orders do the same arithmetic, so results are identical bits.
template <class T, bool PEEL = true>
struct X3TileRunner {
    static constexpr int BM = T::BM, BN = T::BN, BK = T::BK;

    // All threads of the workgroup must call this together.  With `sk`, only
    // k in [split * kchunk, (split + 1) * kchunk) is accumulated and the
    // partials are reduced across splits (common.hpp); `slot` (unique per
    // split tile of the launch, below the workspace's tile count) picks the
    // tile's partials and arrival counter.
    __device__ static void run(const Problem& P, int tm, int tn, char* smem,
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
    __device__ static void run_piece(const Problem& P, int tm, int tn, char* smem,
                                     const SplitK* sk, const KPiece pc) {
        const int t = threadIdx.x;
        const int lane = t & 63, wave = t >> 6;
        const int wm = wave % T::WM, wn = wave / T::WM;
        const int m0 = tm * BM, n0 = tn * BN;
        const int M = P.m, N = P.n, K = P.k;

        // A: thread owns 4x4 (m x k) sub-blocks at k = 4*a_k4, m = 4*(a_m4 + r*A_MSTEP).
        // Loads go through buffer resources rebased per k-tile (see
        // tile_f32.hpp): out-of-extent reads are 0, rows >= M / columns >= N
        // never reach a stored C element, k >= K is zeroed explicitly.
        const int a_k4 = t % (BK / 4);
        const int a_m4 = t / (BK / 4);
        // B: float4 along k, BK/4 lanes per column; each 32-lane group (one
        // 256-byte pass of 8-byte LDS stores) writes rows spaced so that the
        // 2*BK+16-byte row stride puts them on disjoint banks.
        constexpr int S = BK / 4, R = 32 / S, SP = 16 / R;
        static_assert(32 % S == 0 && 16 % R == 0, "B store mapping");
        const int b_k = 4 * (t % S);
        const int b_q = t / 32;
        const int b_n = (b_q % SP) + SP * ((t / S) % R) + 16 * (b_q / SP);
        int a_off[T::A_SUB], b_off[T::B_VEC];
#pragma unroll
        for (int r = 0; r < T::A_SUB; ++r)
            a_off[r] = 4 * (4 * (a_m4 + r * T::A_MSTEP) + 4 * a_k4 * P.lda);
#pragma unroll
        for (int r = 0; r < T::B_VEC; ++r) b_off[r] = 4 * (b_k + (b_n + r * T::B_NSTEP) * P.ldb);
        const long long a_ext = extent(M, K, P.lda), b_ext = extent(K, N, P.ldb);

        // Raw loads land in ra/rb; k >= K masking and the split happen in
        // sstore, after the MFMAs of the current tile.
        f32x4 ra[T::A_SUB][4], rb[T::B_VEC];
        // B's loads first, then A's (see the prologue).
        auto gload = [&](int k0) {
            const long long a_base = m0 + (long long)k0 * P.lda;
            const long long b_base = k0 + (long long)n0 * P.ldb;
            const auto ar = make_rsrc(P.A + a_base, 4 * (a_ext - a_base));
            const auto br = make_rsrc(P.B + b_base, 4 * (b_ext - b_base));
#pragma unroll
            for (int r = 0; r < T::B_VEC; ++r) rb[r] = bload4(br, b_off[r]);
#pragma unroll
            for (int r = 0; r < T::A_SUB; ++r)
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) ra[r][kk] = bload4(ar, a_off[r] + 4 * kk * P.lda);
        };
        // MASK: zero k >= K (only the last k-tile can reach past K).
        auto sstore_impl = [&](char* stage, int k0, auto mask) {
            constexpr bool MASK = decltype(mask)::value;
            char* Ahi = stage;
            char* Alo = stage + T::A_BYTES;
            char* Bhi = stage + 2 * T::A_BYTES;
            char* Blo = Bhi + T::B_BYTES;
            const int ka = K - (k0 + 4 * a_k4);  // valid k in this A sub-block
#pragma unroll
            for (int r = 0; r < T::A_SUB; ++r)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    // row 4*m4 + i gets k = 4*a_k4 .. +3
                    f32x4 col;
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) col[kk] = !MASK || kk < ka ? ra[r][kk][i] : 0.f;
                    bf16x4_t hi, lo;
                    split4(col, hi, lo);
                    const int off = (4 * (a_m4 + r * T::A_MSTEP) + i) * T::RS + 8 * a_k4;
                    *reinterpret_cast<bf16x4_t*>(Ahi + off) = hi;
                    *reinterpret_cast<bf16x4_t*>(Alo + off) = lo;
                }
            const int kb = K - (k0 + b_k);
#pragma unroll
            for (int r = 0; r < T::B_VEC; ++r) {
                f32x4 v;
#pragma unroll
                for (int e = 0; e < 4; ++e) v[e] = !MASK || e < kb ? rb[r][e] : 0.f;
                bf16x4_t hi, lo;
                split4(v, hi, lo);
                const int off = (b_n + r * T::B_NSTEP) * T::RS + 2 * b_k;
                *reinterpret_cast<bf16x4_t*>(Bhi + off) = hi;
                *reinterpret_cast<bf16x4_t*>(Blo + off) = lo;
            }
        };
        auto sstore = [&](char* stage, int k0) {
            if (k0 + BK <= K) sstore_impl(stage, k0, std::false_type{});
            else sstore_impl(stage, k0, std::true_type{});
        };
        // One piece of a full k-tile's staging (the same values and
        // addresses as sstore_impl without the mask): pieces [0, 4 * A_SUB)
        // are the four rows of each A sub-block, the rest B's float4s.
        auto sstore_piece = [&](char* stage, int c) {
            if (c < 4 * T::A_SUB) {
                const int r = c / 4, i = c % 4;
                const f32x4 col = {ra[r][0][i], ra[r][1][i], ra[r][2][i], ra[r][3][i]};
                bf16x4_t hi, lo;
                split4(col, hi, lo);
                const int off = (4 * (a_m4 + r * T::A_MSTEP) + i) * T::RS + 8 * a_k4;
                *reinterpret_cast<bf16x4_t*>(stage + off) = hi;
                *reinterpret_cast<bf16x4_t*>(stage + T::A_BYTES + off) = lo;
            } else {
                const int r = c - 4 * T::A_SUB;
                bf16x4_t hi, lo;
                split4(rb[r], hi, lo);
                const int off = (b_n + r * T::B_NSTEP) * T::RS + 2 * b_k;
                *reinterpret_cast<bf16x4_t*>(stage + 2 * T::A_BYTES + off) = hi;
                *reinterpret_cast<bf16x4_t*>(stage + 2 * T::A_BYTES + T::B_BYTES + off) = lo;
            }
        };

        using Acc = typename std::conditional<T::MF == 16, f32x4, f32x16>::type;
        constexpr int NCRS = T::DUAL ? 1 : 0;
        Acc acc[T::MB][T::NB], crs[NCRS ? T::MB : 1][NCRS ? T::NB : 1];
#pragma unroll
        for (int i = 0; i < T::MB; ++i)
#pragma unroll
            for (int j = 0; j < T::NB; ++j) {
                acc[i][j] = Acc{};
                if constexpr (T::DUAL) crs[i][j] = Acc{};
            }

        // Lane layout of one MFMA operand: row (lane % MF) supplies 8
        // consecutive k at kk + 8 * (lane / MF).
        const int lrow = lane % T::MF, lk = 8 * (lane / T::MF);

        const int kbeg = pc.kbeg;
        const int kend = pc.kend;
        const int ktiles = kend > kbeg ? div_up(kend - kbeg, BK) : 0;
        // SCHED != 0: iterations 0 .. nfast2-1 run interleaved_iteration,
        // which needs k-tile t + 1 full and t + 2 present.
        const int kfull = K - kbeg >= BK ? (K - kbeg) / BK : 0;  // full k-tiles from kbeg
        const int nfast2 = T::SCHED != 0 && PEEL ? max(0, min(kfull - 1, ktiles - 2)) : 0;
        if (ktiles > 0) {
            gload(kbeg);
            sstore(smem, kbeg);
            // k-tile 1 for the first interleaved iteration.  gload issues B,
            // then A, as interleaved_iteration does, so the counts of loads in
            // flight match on entry to its loop from here and from its own
            // last region (the waits at its head take the stricter of the two).
            if (nfast2 > 0) gload(kbeg + BK);
            __syncthreads();
        }
        // A full k-tile t with its loads one k-tile ahead (SCHED != 0).  On
        // entry the staging registers hold k-tile t + 1 (loaded during
        // iteration t - 1); the wave walks its MFMA blocks row by row in
        // regions fenced by sched_barrier.  Region i reads the A operands of
        // row i + 1 and runs row i's MFMAs (each product over all columns
        // before the next, so consecutive MFMAs are independent; per
        // accumulator the order stays al*bh, ah*bl, ah*bh).  Regions 0 ..
        // B_VEC-1 also split B float4 i of k-tile t + 1 into the other LDS
        // stage and at once reload its registers with k-tile t + 2; the next
        // two regions split the A sub-block (it takes four loads, so its
        // registers free up only after its last row), and the regions after
        // them reload it.  Each load so has most of a k-tile to land, and
        // the split runs in the shadow of the matrix cores.
        auto interleaved_iteration = [&](int kt) {
            if constexpr (T::SCHED != 0) {
                constexpr int SUB = T::MB >= 8 ? 1 : 8 / T::MB;
                constexpr int NREG = T::MB * SUB;
                static_assert(T::MF == 16 && BK == T::KSTEP && T::A_SUB == 1 &&
                                  T::NB % SUB == 0 && NREG >= T::B_VEC + 3,
                              "interleaved body: one k-step, B pieces, two A regions, A reloads");
                const char* stage = smem + (kt & 1) * T::STAGE;
                char* next_stage = smem + ((kt + 1) & 1) * T::STAGE;
                const char* Ahi = stage;
                const char* Alo = stage + T::A_BYTES;
                const char* Bhi = stage + 2 * T::A_BYTES;
                const char* Blo = Bhi + T::B_BYTES;
                const int k2 = kbeg + (kt + 2) * BK;
                const long long a_base = m0 + (long long)k2 * P.lda;
                const long long b_base = k2 + (long long)n0 * P.ldb;
                const auto ar = make_rsrc(P.A + a_base, 4 * (a_ext - a_base));
                const auto br = make_rsrc(P.B + b_base, 4 * (b_ext - b_base));
                bf16x8_t ah[T::MB], al[T::MB], bh[T::NB], bl[T::NB];
#pragma unroll
                for (int j = 0; j < T::NB; ++j) {
                    const int off = (wn * T::TN + 16 * j + lrow) * T::RS + 2 * lk;
                    bh[j] = *reinterpret_cast<const bf16x8_t*>(Bhi + off);
                    bl[j] = *reinterpret_cast<const bf16x8_t*>(Blo + off);
                }
                auto read_a = [&](int i) {
                    const int off = (wm * T::TM + 16 * i + lrow) * T::RS + 2 * lk;
                    ah[i] = *reinterpret_cast<const bf16x8_t*>(Ahi + off);
                    al[i] = *reinterpret_cast<const bf16x8_t*>(Alo + off);
                };
                read_a(0);
                // NREG regions: row i of blocks in SUB parts of NB / SUB
                // columns each (SUB = 2 for a four-row wave tile, so there are
                // eight regions either way).
#pragma unroll
                for (int r = 0; r < NREG; ++r) {
                    const int i = r / SUB, j0 = (r % SUB) * (T::NB / SUB);
                    if (r % SUB == 0 && i + 1 < T::MB) read_a(i + 1);
#pragma unroll
                    for (int j = j0; j < j0 + T::NB / SUB; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
                    for (int j = j0; j < j0 + T::NB / SUB; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
                    for (int j = j0; j < j0 + T::NB / SUB; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
                    if (r < T::B_VEC) {
                        sstore_piece(next_stage, 4 + r);
                        rb[r] = bload4(br, b_off[r]);
                    }
                    if (r >= T::B_VEC && r < T::B_VEC + 2) {
                        sstore_piece(next_stage, 2 * (r - T::B_VEC));
                        sstore_piece(next_stage, 2 * (r - T::B_VEC) + 1);
                    }
                    // One A reload per region from the one after the last A
                    // row on (the rest in the last region), so no region
                    // queues many loads at once.
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk)
                        if (r == min(T::B_VEC + 2 + kk, NREG - 1))
                            ra[0][kk] = bload4(ar, a_off[0] + 4 * kk * P.lda);
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        };
        // One k-tile: MFMAs on the current stage while the next k-tile is
        // loaded, split and stored to the other stage.  NEXT_FULL: a next
        // k-tile exists and lies entirely below K, so the body needs no
        // masking and no branch -- one basic block, which the scheduling
        // hints below can interleave.
        auto iteration = [&](int kt, auto next_full_tag, bool preloaded = false) {
            constexpr bool NEXT_FULL = decltype(next_full_tag)::value;
            const char* stage = smem + (kt & 1) * T::STAGE;
            char* next_stage = smem + ((kt + 1) & 1) * T::STAGE;
            const char* Ahi = stage;
            const char* Alo = stage + T::A_BYTES;
            const char* Bhi = stage + 2 * T::A_BYTES;
            const char* Blo = Bhi + T::B_BYTES;
            const bool more = NEXT_FULL || kt + 1 < ktiles;
            const int knext = kbeg + (kt + 1) * BK;
            if (more && !preloaded) gload(knext);
            if constexpr (NEXT_FULL) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int kk = 0; kk < BK; kk += T::KSTEP) {
                bf16x8_t ah[T::MB], al[T::MB], bh[T::NB], bl[T::NB];
#pragma unroll
                for (int i = 0; i < T::MB; ++i) {
                    const int off = (wm * T::TM + T::MF * i + lrow) * T::RS + 2 * (kk + lk);
                    ah[i] = *reinterpret_cast<const bf16x8_t*>(Ahi + off);
                    al[i] = *reinterpret_cast<const bf16x8_t*>(Alo + off);
                }
#pragma unroll
                for (int j = 0; j < T::NB; ++j) {
                    const int off = (wn * T::TN + T::MF * j + lrow) * T::RS + 2 * (kk + lk);
                    bh[j] = *reinterpret_cast<const bf16x8_t*>(Bhi + off);
                    bl[j] = *reinterpret_cast<const bf16x8_t*>(Blo + off);
                }
#pragma unroll
                for (int i = 0; i < T::MB; ++i)
#pragma unroll
                    for (int j = 0; j < T::NB; ++j) {
                        if constexpr (T::MF == 16) {
                            if constexpr (T::DUAL) {
                                crs[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], crs[i][j], 0, 0, 0);
                                crs[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], crs[i][j], 0, 0, 0);
                            } else {
                                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                            }
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
                        } else {
                            if constexpr (T::DUAL) {
                                crs[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], crs[i][j], 0, 0, 0);
                                crs[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], crs[i][j], 0, 0, 0);
                            } else {
                                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                            }
                            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
                        }
                    }
                // Stage the next k-tile (wait on its loads, split, LDS stores)
                // behind the first k-step's MFMAs, or after the last one.
                constexpr int kStoreAt = T::MID ? 0 : BK - T::KSTEP;
                if (kk == kStoreAt) {
                    if constexpr (NEXT_FULL) {
                        // Keep the staging code a region of its own (as a
                        // branch would): letting the scheduler mix it into
                        // the MFMA section raises register pressure past the
                        // budget.
                        if constexpr (T::SCHED == 0) __builtin_amdgcn_sched_barrier(0);
                        sstore_impl(next_stage, knext, std::false_type{});
                        if constexpr (T::SCHED == 0) __builtin_amdgcn_sched_barrier(0);
                    } else if (more) {
                        sstore(next_stage, knext);
                    }
                }
            }
        };
        // k-tiles whose successor is full run the branch-free body; the
        // last one or two run the general body.  (Two loops in sequence, not
        // one loop with two bodies, so the accumulators never merge from two
        // paths -- that doubles their registers.)
        const int nfast = min(ktiles, kfull) - 1;  // iterations with a full successor
        int kt = 0;
        if constexpr (T::SCHED != 0 && PEEL) {
            for (; kt < nfast2; ++kt) {
                interleaved_iteration(kt);
                __syncthreads();
            }
            // k-tile kt + 1 is already in the staging registers.
            if (nfast2 > 0 && kt < ktiles) {
                iteration(kt, std::false_type{}, true);
                __syncthreads();
                ++kt;
            }
        } else if constexpr (PEEL) {
            for (; kt < nfast; ++kt) {
                iteration(kt, std::true_type{});
                __syncthreads();
            }
        }
        for (; kt < ktiles; ++kt) {
            iteration(kt, std::false_type{});
            __syncthreads();
        }

        if (sk && pc.n > 1) {
            // Fold the cross-term accumulators in, dump this piece's sums in
            // lane order, and let the last workgroup of the tile add all
            // pieces in k order.
            if constexpr (T::DUAL) {
#pragma unroll
                for (int i = 0; i < T::MB; ++i)
#pragma unroll
                    for (int j = 0; j < T::NB; ++j) {
                        acc[i][j] += crs[i][j];
                        crs[i][j] = Acc{};
                    }
            }
            constexpr int NV = sizeof(Acc) / 16;  // float4 pieces per accumulator
            constexpr int NACC = T::MB * T::NB * NV;
            constexpr int SPLIT_BYTES = NACC * T::NT * 16;
            const int tile = pc.region;
            const __amdgpu_buffer_rsrc_t part =
                make_rsrc(sk->partials + (size_t)tile * pc.stride * NACC * T::NT * 4,
                          (long long)pc.stride * SPLIT_BYTES);
            auto piece = [&](int i, int j, int v) -> f32x4 {
                return f32x4{acc[i][j][4 * v], acc[i][j][4 * v + 1], acc[i][j][4 * v + 2],
                             acc[i][j][4 * v + 3]};
            };
            auto off = [&](int i, int j, int v) { return (((i * T::NB + j) * NV + v) * T::NT + t) * 16; };
#pragma unroll
            for (int i = 0; i < T::MB; ++i)
#pragma unroll
                for (int j = 0; j < T::NB; ++j)
#pragma unroll
                    for (int v = 0; v < NV; ++v)
                        store_agent(part, pc.q * SPLIT_BYTES + off(i, j, v), piece(i, j, v));
            if (!splitk_arrive(*sk, tile, reinterpret_cast<int*>(smem), pc.n)) return;
            // This workgroup's own partial is re-read like the others
            // (bit-identical to its registers), so the sum reuses the
            // accumulators' registers.
#pragma unroll
            for (int i = 0; i < T::MB; ++i)
#pragma unroll
                for (int j = 0; j < T::NB; ++j)
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        const f32x4 x = load_agent(part, off(i, j, v));
#pragma unroll
                        for (int r = 0; r < 4; ++r) acc[i][j][4 * v + r] = x[r];
                    }
            for (int q = 1; q < pc.n; ++q) {
#pragma unroll
                for (int i = 0; i < T::MB; ++i)
#pragma unroll
                    for (int j = 0; j < T::NB; ++j)
#pragma unroll
                        for (int v = 0; v < NV; ++v) {
                            const f32x4 x = load_agent(part, q * SPLIT_BYTES + off(i, j, v));
#pragma unroll
                            for (int r = 0; r < 4; ++r) acc[i][j][4 * v + r] += x[r];
                        }
            }
        }

        // Epilogue through LDS (epilogue_block): whole 16-byte pieces down
        // each column of C.  16x16 block: lane holds rows 4*(lane/16) + 0..3
        // of column lane % 16; 32x32 block: register group q (4 registers)
        // holds rows 8q + 4*(lane/32) + 0..3 of column lane % 32.
        float* elds = reinterpret_cast<float*>(smem) + kEpilogueLdsSkip +
                      wave * T::MF * (T::TM + 4);
#pragma unroll
        for (int j = 0; j < T::NB; ++j)
            epilogue_block<T::TM, T::MF, T::MB>(
                elds, P, m0 + wm * T::TM, n0 + wn * T::TN + T::MF * j,
                [&](int i, int q) {
                    f32x4 s;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        s[r] = acc[i][j][4 * q + r];
                        if constexpr (T::DUAL) s[r] += crs[i][j][4 * q + r];
                    }
                    return s;
                },
                [&](int, const float* cc, int valid) { return load4(cc, valid); });
    }
};

// Config ids of the BF16x3 form (the order is the ABI's config id).
using X3_128x128 = X3Tile<128, 128, 32, 2, 2>;
using X3_64x64 = X3Tile<64, 64, 64, 2, 2>;
// 16x16x32 MFMAs on the 256x256 tile: 389 -> 413 TFLOP/s at 8192^3 and
// 385 -> 398 at 16384^3 over 32x32x16 with scheduling hints.
// The next k-tile's split interleaved with the MFMAs and its loads issued a
// k-tile ahead (SCHED 1): 8192^3 404 -> 439, 16384^3 377 -> 427, 3072^3
// 251 -> 292 TFLOP/s over the split staged after the MFMAs, same bits.
using X3_256x256_w8 = X3Tile<256, 256, 32, 2, 4, 16, false, false, 0, 1>;
using X3_256x128_w4 = X3Tile<256, 128, 32, 2, 2, 32, false>;
using X3_128x128_m32 = X3Tile<128, 128, 32, 2, 2, 32, true>;
using X3_256x128_k16 = X3Tile<256, 128, 16, 2, 2, 32, false, true>;
// 16x16x32 MFMAs: the grouped BF16x3 benchmark 97 -> 102 TFLOP/s, config
// 10 at 2048^3 246 -> 263 TFLOP/s over 32x32x16.
// Interleaved k loop as config 2's (SCHED 1): 2048^3 200 -> 275, 1536^3 112
// -> 153, 8192^3 316 -> 338 TFLOP/s, same bits.
using X3_128x128_2cu = X3Tile<128, 128, 32, 2, 2, 16, false, false, 2, 1>;

template <class F>
__host__ __device__ inline bool dispatch_x3(int id, F&& f) {
    switch (id) {
        case 0: f(TypeTag<X3_128x128>{}); return true;
        case 1: f(TypeTag<X3_64x64>{}); return true;
        case 2: f(TypeTag<X3_256x256_w8>{}); return true;
        case 3: f(TypeTag<X3_256x128_w4>{}); return true;
        case 4: f(TypeTag<X3_128x128_m32>{}); return true;
        case 5: f(TypeTag<X3_256x128_k16>{}); return true;
        case 8: f(TypeTag<X3_128x128_2cu>{}); return true;
        default: return false;
    }
}

}  // namespace sgemm_amd
