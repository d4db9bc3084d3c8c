/*
 * oracle.c -- CPU oracle for the column-major SGEMM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py load liboracle.so, and only as the checker (or, for the CPU
 * baseline, as the timed host restatement).
 *
 * What is here, and which reference code each part restates
 * (paths relative to the szrlee/sgemm.cu tree):
 *
 *   orc_splitmix64 / orc_fill_*   Seeded input generation.  The reference seeds
 *                                 rand() by wall-clock time (test.cu:19,
 *                                 common/helper_matrix.h:20-23); this oracle
 *                                 uses SplitMix64 so inputs are reproducible.
 *   orc_dgemm_cm                  FP64 textbook GEMM: ground truth for every
 *                                 form.  Column-major BLAS "NN" semantics; the
 *                                 reference's row-major sgemm (src/sgemm.cuh:19-33)
 *                                 maps onto it by swapping operands.
 *   orc_dgemm_entries_cm          The same, for a list of (i, j) entries
 *                                 (full-size checks sample entries).
 *   orc_sgemm_basic_rm            FP32 restatement of kernel_basic
 *                                 (src/kernels/basic.cuh:15-25): sequential-k
 *                                 FMA accumulation, then *alpha, then +C*beta
 *                                 (always reads C).  This is also the timed CPU
 *                                 baseline (OpenMP over output rows).
 *   orc_bf16_rne / orc_bf16x3_split  IEEE round-to-nearest-even FP32 -> BF16 and
 *                                 the hi/lo split used by the BF16x3 form.
 *   orc_check                     Componentwise error metric and the k-scaled
 *                                 tolerance  tol_ij = 2*gamma_{k+2}*scale_ij,
 *                                 scale = |alpha|(|A||B|) + |beta||C0|; NaN/Inf
 *                                 in the candidate always fail, as in
 *                                 common/helper_matrix.h:63-78.
 *
 * Pinning: every function above is checked in tests/test_oracle.py against
 * values that do not come from this file -- the published SplitMix64 stream,
 * exact rational arithmetic (Python fractions) on golden fixtures, numpy's
 * float64 matmul, closed forms (ones / identity / permutation matrices) and
 * hand-computed BF16 rounding cases.
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define ORC_GOLDEN_GAMMA 0x9E3779B97F4A7C15ULL

static inline uint64_t orc_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* One SplitMix64 step (Steele, Lea, Flood 2014; Vigna's splitmix64.c). */
uint64_t orc_splitmix64(uint64_t* state) {
    *state += ORC_GOLDEN_GAMMA;
    return orc_mix64(*state);
}

/* The i-th output (0-based) of the stream seeded with `seed`.  Counter form of
 * the same generator, so the fills below parallelise and stay identical to a
 * sequential walk. */
uint64_t orc_splitmix64_at(uint64_t seed, uint64_t i) {
    return orc_mix64(seed + (i + 1) * ORC_GOLDEN_GAMMA);
}

/* u in [0, 1) with 24 random bits: exact in FP32. */
static inline double orc_u24(uint64_t z) { return (double)(z >> 40) * (1.0 / 16777216.0); }

/* Uniform in [-1, 1): 2u - 1, exact in FP32. */
void orc_fill_uniform(float* x, size_t n, uint64_t seed) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) x[i] = (float)(2.0 * orc_u24(orc_splitmix64_at(seed, i)) - 1.0);
}

/* Wide dynamic range: sign * 2^e, e uniform in [-8, 8) (24-bit u), random sign
 * from bit 39 of the same draw.  exp2 is evaluated in double, then rounded. */
void orc_fill_wide(float* x, size_t n, uint64_t seed) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) {
        uint64_t z = orc_splitmix64_at(seed, i);
        double e = -8.0 + 16.0 * orc_u24(z);
        double v = exp2(e);
        x[i] = (float)(((z >> 39) & 1) ? -v : v);
    }
}

/* Integers uniform in [lo, hi] (inclusive), from the high 32 bits. */
void orc_fill_int(float* x, size_t n, uint64_t seed, int lo, int hi) {
    uint64_t span = (uint64_t)((int64_t)hi - (int64_t)lo + 1);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) {
        uint64_t z = orc_splitmix64_at(seed, i);
        x[i] = (float)(lo + (int64_t)((z >> 32) % span));
    }
}

/* FNV-1a over the bytes of a buffer: logged input hash. */
uint64_t orc_fnv1a64(const void* data, size_t nbytes) {
    const unsigned char* p = (const unsigned char*)data;
    uint64_t h = 0xcbf29ce484222325ULL;
    for (size_t i = 0; i < nbytes; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ULL;
    }
    return h;
}

/*
 * FP64 ground truth, column-major:
 *   out[i + j*ldo]   = alpha * sum_p A[i + p*lda] * B[p + j*ldb] + beta * C0[i + j*ldc]
 *   scale[i + j*ldo] = |alpha| * sum_p |A||B|            + |beta| * |C0|
 * beta == 0 means C0 is not read (BLAS convention; the hot kernels of the
 * reference skip the read too, src/kernels/128x128x8.cuh:364-380).
 * `scale` may be NULL.  k == 0 gives out = beta*C0.
 */
void orc_dgemm_cm(int m, int n, int k, double alpha, const float* A, int lda, const float* B,
                  int ldb, double beta, const float* C0, int ldc, double* out, double* scale,
                  int ldo) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j < n; ++j) {
        double* o = out + (size_t)j * ldo;
        double* s = scale ? scale + (size_t)j * ldo : NULL;
        for (int i = 0; i < m; ++i) {
            o[i] = 0.0;
            if (s) s[i] = 0.0;
        }
        for (int p = 0; p < k; ++p) {
            const float* a = A + (size_t)p * lda;
            double b = (double)B[p + (size_t)j * ldb];
            double ab = fabs(b);
            for (int i = 0; i < m; ++i) o[i] += (double)a[i] * b;
            if (s)
                for (int i = 0; i < m; ++i) s[i] += fabs((double)a[i]) * ab;
        }
        const float* c = (beta != 0.0 && C0) ? C0 + (size_t)j * ldc : NULL;
        for (int i = 0; i < m; ++i) {
            double v = alpha * o[i];
            if (c) v += beta * (double)c[i];
            o[i] = v;
            if (s) s[i] = fabs(alpha) * s[i] + (c ? fabs(beta) * fabs((double)c[i]) : 0.0);
        }
    }
}

/* Same as orc_dgemm_cm for `count` entries (ii[t], jj[t]). */
void orc_dgemm_entries_cm(int k, double alpha, const float* A, int lda, const float* B, int ldb,
                          double beta, const float* C0, int ldc, int count, const int* ii,
                          const int* jj, double* out, double* scale) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int t = 0; t < count; ++t) {
        int i = ii[t], j = jj[t];
        double acc = 0.0, sacc = 0.0;
        for (int p = 0; p < k; ++p) {
            double a = (double)A[i + (size_t)p * lda];
            double b = (double)B[p + (size_t)j * ldb];
            acc += a * b;
            sacc += fabs(a) * fabs(b);
        }
        double v = alpha * acc, s = fabs(alpha) * sacc;
        if (beta != 0.0 && C0) {
            double c = (double)C0[i + (size_t)j * ldc];
            v += beta * c;
            s += fabs(beta) * fabs(c);
        }
        out[t] = v;
        if (scale) scale[t] = s;
    }
}

/*
 * FP32 restatement of kernel_basic (src/kernels/basic.cuh:15-25), row-major:
 *   acc = 0; for p in 0..k-1: acc = fma(A[y*lda+p], B[x+p*ldb], acc)
 *   acc *= alpha; acc = fma(C[x+y*ldc], beta, acc)     (C is always read)
 * nvcc contracts `acc += a*b` and `acc += c*beta` into FMAs by default; fmaf
 * states that explicitly.  Loop order y-p-x keeps each output's p-order while
 * the x loop vectorises.  `acc` is a per-thread row buffer of n floats.
 */
#define ORC_BASIC_BODY                                                              \
    _Pragma("omp parallel")                                                         \
    {                                                                               \
        float* acc = (float*)malloc((size_t)(n > 0 ? n : 1) * sizeof(float));       \
        _Pragma("omp for schedule(static)")                                         \
        for (int y = 0; y < m; ++y) {                                               \
            for (int x = 0; x < n; ++x) acc[x] = 0.0f;                              \
            for (int p = 0; p < k; ++p) {                                           \
                float a = A[(size_t)y * lda + p];                                   \
                const float* b = B + (size_t)p * ldb;                               \
                for (int x = 0; x < n; ++x) acc[x] = fmaf(a, b[x], acc[x]);         \
            }                                                                       \
            float* c = C + (size_t)y * ldc;                                         \
            for (int x = 0; x < n; ++x) {                                           \
                float v = acc[x] * alpha;                                           \
                c[x] = fmaf(c[x], beta, v);                                         \
            }                                                                       \
        }                                                                           \
        free(acc);                                                                  \
    }

__attribute__((target("avx2,fma"))) static void orc_sgemm_basic_rm_fma(
    int m, int n, int k, float alpha, const float* A, int lda, const float* B, int ldb,
    float beta, float* C, int ldc) {
    ORC_BASIC_BODY
}

static void orc_sgemm_basic_rm_generic(int m, int n, int k, float alpha, const float* A, int lda,
                                       const float* B, int ldb, float beta, float* C, int ldc) {
    ORC_BASIC_BODY
}

void orc_sgemm_basic_rm(int m, int n, int k, float alpha, const float* A, int lda,
                        const float* B, int ldb, float beta, float* C, int ldc) {
    __builtin_cpu_init();
    if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma"))
        orc_sgemm_basic_rm_fma(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    else
        orc_sgemm_basic_rm_generic(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

/* Column-major view of the same computation: C^T = B^T A^T, so a column-major
 * (m, n, k) product is the row-major (n, m, k) product with A and B swapped
 * (the swap the reference's own cuBLAS call makes, benchmark.cu:150-164). */
void orc_sgemm_basic_cm(int m, int n, int k, float alpha, const float* A, int lda,
                        const float* B, int ldb, float beta, float* C, int ldc) {
    orc_sgemm_basic_rm(n, m, k, alpha, B, ldb, A, lda, beta, C, ldc);
}

/* IEEE-754 binary32 -> bfloat16, round to nearest, ties to even.  NaN stays a
 * (quiet) NaN with its sign; overflow rounds to +-Inf as the definition says. */
uint16_t orc_bf16_rne(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x0040u);
    uint32_t bias = 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)((u + bias) >> 16);
}

float orc_bf16_to_f32(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

/* hi = bf16(clamp(x, -BF16_MAX, BF16_MAX)), lo = bf16(x - float(hi)).
 * BF16_MAX = 0x7f7f (3.38953139e38) is the largest finite bf16; the clamp is
 * the identity for |x| <= BF16_MAX, so there hi = bf16(x).  Beyond it hi
 * stays finite: a finite x keeps an exact residual x - hi (Sterbenz), and
 * x = +-Inf puts the infinity into lo.  For NaN x only lo = NaN is specified
 * (the GPU's clamp of a NaN is not pinned); here hi is NaN too. */
void orc_bf16x3_split(const float* x, size_t n, uint16_t* hi, uint16_t* lo) {
    const float bmax = orc_bf16_to_f32(0x7f7f);
    for (size_t i = 0; i < n; ++i) {
        float c = x[i];
        if (c > bmax) c = bmax;
        if (c < -bmax) c = -bmax;
        uint16_t h = orc_bf16_rne(c);
        float r = x[i] - orc_bf16_to_f32(h);
        hi[i] = h;
        lo[i] = orc_bf16_rne(r);
    }
}

/* Result of orc_check. */
typedef struct {
    double max_err;      /* max_ij |C - ref| / scale      (componentwise) */
    double mean_err;     /* mean of the same                               */
    double max_rel;      /* max_ij |C - ref| / |ref| over ref != 0          */
    double mean_rel;     /* mean of the same                               */
    double max_ratio;    /* max_ij |C - ref| / tol_ij  (pass iff <= 1)     */
    long long n_fail;    /* entries with |C - ref| > tol, incl. NaN/Inf    */
    long long n_nan;
    long long n_inf;
    long long n_checked;
} orc_stats;

/* gamma_k = k*u / (1 - k*u), u = 2^-24 (FP32 unit roundoff). */
double orc_gamma(double k) {
    double ku = k * 5.9604644775390625e-08;
    return ku / (1.0 - ku);
}

/*
 * Compare a candidate (float, column-major, ldc) against ref/scale (double,
 * ldo) with tol_ij = tol_mult * scale_ij.  A zero tolerance demands equality.
 */
void orc_check(int m, int n, const float* C, int ldc, const double* ref, const double* scale,
               int ldo, double tol_mult, orc_stats* st) {
    double max_err = 0, sum_err = 0, max_rel = 0, sum_rel = 0, max_ratio = 0;
    long long n_fail = 0, n_nan = 0, n_inf = 0, n_rel = 0;
#pragma omp parallel for schedule(static) reduction(max : max_err, max_rel, max_ratio) \
    reduction(+ : sum_err, sum_rel, n_fail, n_nan, n_inf, n_rel)
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < m; ++i) {
            double c = (double)C[i + (size_t)j * ldc];
            double r = ref[i + (size_t)j * ldo];
            double s = scale[i + (size_t)j * ldo];
            if (isnan(c)) { n_nan++; n_fail++; max_ratio = INFINITY; continue; }
            if (isinf(c)) { n_inf++; n_fail++; max_ratio = INFINITY; continue; }
            double d = fabs(c - r);
            double e = s > 0 ? d / s : (d > 0 ? INFINITY : 0.0);
            double tol = tol_mult * s;
            double ratio = tol > 0 ? d / tol : (d > 0 ? INFINITY : 0.0);
            if (e > max_err) max_err = e;
            sum_err += e;
            if (ratio > max_ratio) max_ratio = ratio;
            if (d > tol) n_fail++;
            if (r != 0.0) {
                double rel = d / fabs(r);
                if (rel > max_rel) max_rel = rel;
                sum_rel += rel;
                n_rel++;
            }
        }
    }
    long long cnt = (long long)m * n;
    st->max_err = max_err;
    st->mean_err = cnt ? sum_err / (double)cnt : 0.0;
    st->max_rel = max_rel;
    st->mean_rel = n_rel ? sum_rel / (double)n_rel : 0.0;
    st->max_ratio = max_ratio;
    st->n_fail = n_fail;
    st->n_nan = n_nan;
    st->n_inf = n_inf;
    st->n_checked = cnt;
}

int orc_max_threads(void) { return omp_get_max_threads(); }
