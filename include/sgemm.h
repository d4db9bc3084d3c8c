/*
 * sgemm.h -- C ABI of the MI355X (gfx950) SGEMM library libsgemm_amd.so.
 *
 * Every entry point is extern "C" with plain pointers and sizes.  All matrices
 * are COLUMN-MAJOR (BLAS "NN"): element (i, j) of an (rows x cols) operand with
 * leading dimension ld is at ptr[i + j*ld].  A, B, C are device pointers; alpha
 * and beta are host pointers read before the call returns.
 *
 * Reference interface replaced: szrlee/sgemm.cu exposes one C++ function,
 *   void sgemm(int m, int n, int k, const float* alpha, float* A, int lda,
 *              float* B, int ldb, const float* beta, float* C, int ldc)
 * defined in src/sgemm.cuh:19-90, plus the naive test kernel launcher
 * sgemm_basic (src/sgemm.cuh:92-113).  The reference's operands are row-major
 * (src/sgemm.cuh:31-33); this library follows the column-major convention and
 * offers sgemm_rm for callers that want the reference's row-major meaning.
 */
#ifndef SGEMM_AMD_SGEMM_H
#define SGEMM_AMD_SGEMM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A hipStream_t.  NULL is the null stream, which the reference launches on. */
typedef void* sgemm_stream_t;

/* Status codes returned by the *_ex entry points and recorded for
 * sgemm_get_last_error() by the void entry points. */
enum {
    SGEMM_STATUS_SUCCESS = 0,
    SGEMM_STATUS_INVALID_VALUE = 1,   /* negative size, bad leading dimension, NULL operand */
    SGEMM_STATUS_NOT_SUPPORTED = 2,   /* unknown form / config id */
    SGEMM_STATUS_HIP_ERROR = 3,       /* a HIP runtime call or kernel launch failed */
    SGEMM_STATUS_ALLOC_FAILED = 4,    /* device or pinned host allocation failed */
    SGEMM_STATUS_BUDGET_EXCEEDED = 5, /* a request can never fit the server's HBM budget */
    SGEMM_STATUS_NOT_READY = 6,       /* a polled request is still in flight */
    SGEMM_STATUS_NOT_FOUND = 7,       /* unknown handle or request id */
    SGEMM_STATUS_IO_ERROR = 8,        /* autotune table could not be read or written */
    SGEMM_STATUS_WRONG_STATE = 9      /* e.g. replay of a plan that was never captured */
};

/* Arithmetic forms. */
enum {
    SGEMM_FORM_F32 = 0,    /* plain FP32: FP32 products, FP32 accumulation */
    SGEMM_FORM_BF16X3 = 1  /* A*B ~ hi*hi + hi*lo + lo*hi on BF16 MFMA, FP32 accumulation */
};

/* ------------------------------------------------------------------------ */
/* Drop-in entry points (reference: src/sgemm.cuh:19-30 and :92-113).        */
/* ------------------------------------------------------------------------ */

/* C := alpha*A*B + beta*C, column-major, FP32, on the null stream,
 * asynchronous.  A is m x k (lda >= max(1,m)), B is k x n (ldb >= max(1,k)),
 * C is m x n (ldc >= max(1,m)).  beta == 0 means C is not read.
 * m == 0 or n == 0 is a no-op; k == 0 gives C := beta*C.  Invalid arguments
 * launch nothing and are reported by sgemm_get_last_error().
 * Replaces sgemm() of src/sgemm.cuh:19-48 (which is row-major and ignores
 * ldb/ldc in its fast kernels; this one honours every leading dimension). */
void sgemm(int m, int n, int k, const float* alpha, float* A, int lda, float* B, int ldb,
           const float* beta, float* C, int ldc);

/* The naive one-thread-per-output kernel, column-major: acc = sum_p A*B in p
 * order with FMA, acc *= alpha, acc = fma(C, beta, acc); C is always read.
 * Replaces sgemm_basic()/kernel_basic of src/sgemm.cuh:92-113 and
 * src/kernels/basic.cuh:1-26 (the reference's own test oracle). */
void sgemm_basic(int m, int n, int k, const float* alpha, float* A, int lda, float* B, int ldb,
                 const float* beta, float* C, int ldc);

/* The reference's row-major meaning of sgemm (src/sgemm.cuh:31-33): A is m x k
 * row-major with lda, B is k x n row-major with ldb, C is m x n row-major with
 * ldc.  Computed as the column-major product C^T = B^T A^T. */
void sgemm_rm(int m, int n, int k, const float* alpha, float* A, int lda, float* B, int ldb,
              const float* beta, float* C, int ldc);

/* Status of the last void entry point called on this host thread (then
 * cleared).  The reference reports nothing (src/sgemm.cuh:19-90). */
int sgemm_get_last_error(void);
const char* sgemm_status_string(int status);

/* ------------------------------------------------------------------------ */
/* Status-returning, stream-aware forms (no reference counterpart; the       */
/* reference's only entry point is sgemm() above).                           */
/* ------------------------------------------------------------------------ */

/* As sgemm(), on `stream`, returning a status.  The config is chosen by the
 * autotune table for (m, n, k, SGEMM_FORM_F32). */
int sgemm_ex(int m, int n, int k, const float* alpha, const float* A, int lda, const float* B,
             int ldb, const float* beta, float* C, int ldc, sgemm_stream_t stream);

/* Split-precision form: each FP32 input x is split on the GPU into
 * hi = bf16_rne(x), lo = bf16_rne(x - hi); A*B is accumulated in FP32 as
 * hi*hi + hi*lo + lo*hi on BF16 MFMA.  Same operand conventions as sgemm_ex. */
int sgemm_bf16x3(int m, int n, int k, const float* alpha, const float* A, int lda,
                 const float* B, int ldb, const float* beta, float* C, int ldc,
                 sgemm_stream_t stream);

/* Explicit form and kernel config (config ids from sgemm_config_count /
 * sgemm_config_name).  Used by the autotuner and by tests. */
int sgemm_run_config(int form, int config_id, int m, int n, int k, const float* alpha,
                     const float* A, int lda, const float* B, int ldb, const float* beta,
                     float* C, int ldc, sgemm_stream_t stream);
int sgemm_config_count(int form);
/* Writes "BMxBNxBK/WMxWN" style text; returns the tile sizes. */
int sgemm_config_info(int form, int config_id, int* bm, int* bn, int* bk, char* name,
                      int name_len);

/* Library version string. */
const char* sgemm_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SGEMM_AMD_SGEMM_H */
