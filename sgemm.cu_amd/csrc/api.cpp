// api.cpp -- the C ABI (include/sgemm.h): argument validation, status
// reporting and kernel-config dispatch for the single-problem entry points.
//
// Reference boundary: sgemm() and sgemm_basic() of src/sgemm.cuh:19-113.  The
// reference validates nothing and returns void; here invalid arguments launch
// nothing and are reported through a status (the *_ex forms) or through
// sgemm_get_last_error() (the void drop-in forms).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>

#include "sgemm.h"
#include "autotune.hpp"
#include "kernels.hpp"

using namespace sgemm_amd;

namespace {

thread_local int g_last_error = SGEMM_STATUS_SUCCESS;
thread_local hipError_t g_last_hip_error = hipSuccess;

int record(int status) {
    g_last_error = status;
    return status;
}

}  // namespace

// Column-major BLAS NN argument rules.  A is m x k, B is k x n, C is m x n.
int sgemm_amd::validate_args(int m, int n, int k, const float* alpha, const void* A, int lda,
                             const void* B, int ldb, const float* beta, const void* C,
                             int ldc) {
    if (m < 0 || n < 0 || k < 0) return SGEMM_STATUS_INVALID_VALUE;
    if (alpha == nullptr || beta == nullptr) return SGEMM_STATUS_INVALID_VALUE;
    if (lda < (m > 1 ? m : 1) || ldb < (k > 1 ? k : 1) || ldc < (m > 1 ? m : 1))
        return SGEMM_STATUS_INVALID_VALUE;
    if (m == 0 || n == 0) return SGEMM_STATUS_SUCCESS;
    if (C == nullptr || !dword_aligned(C)) return SGEMM_STATUS_INVALID_VALUE;
    if (k > 0 && *alpha != 0.f &&
        (A == nullptr || B == nullptr || !dword_aligned(A) || !dword_aligned(B)))
        return SGEMM_STATUS_INVALID_VALUE;
    return SGEMM_STATUS_SUCCESS;
}

namespace {

int validate(int m, int n, int k, const float* alpha, const void* A, int lda, const void* B,
             int ldb, const float* beta, const void* C, int ldc) {
    return validate_args(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

Problem make_problem(int m, int n, int k, float alpha, const float* A, int lda, const float* B,
                     int ldb, float beta, float* C, int ldc) {
    Problem p;
    std::memset(&p, 0, sizeof(p));
    p.A = A;
    p.B = B;
    p.C = C;
    p.m = m;
    p.n = n;
    p.k = k;
    p.lda = lda;
    p.ldb = ldb;
    p.ldc = ldc;
    p.alpha = alpha;
    p.beta = beta;
    return p;
}

int hip_status(hipError_t e) {
    if (e == hipSuccess) return SGEMM_STATUS_SUCCESS;
    g_last_hip_error = e;
    return SGEMM_STATUS_HIP_ERROR;
}

}  // namespace

// Shared body of every single-problem entry point.  config_id < 0 selects the
// config from the autotune table.
int sgemm_amd::run_problem(int form, int config_id, int m, int n, int k, const float* alpha,
                           const float* A, int lda, const float* B, int ldb, const float* beta,
                           float* C, int ldc, hipStream_t stream, const Arena* arena) {
    int st = validate(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    if (st != SGEMM_STATUS_SUCCESS) return st;
    if (form != SGEMM_FORM_F32 && form != SGEMM_FORM_BF16X3) return SGEMM_STATUS_NOT_SUPPORTED;
    if (config_id >= config_count(form)) return SGEMM_STATUS_NOT_SUPPORTED;
    if (m == 0 || n == 0) return SGEMM_STATUS_SUCCESS;
    Problem p = make_problem(m, n, k, *alpha, A, lda, B, ldb, *beta, C, ldc);
    // No product term: C := beta*C (BLAS: A and B are not referenced).
    if (k == 0 || *alpha == 0.f) return hip_status(launch_scale_c(p, stream));
    // Leading dimensions past the tiles' 32-bit buffer offsets: the default
    // selection falls back to the naive FP32 kernel (64-bit addressing); an
    // explicit config cannot run such a problem.
    if (lda > kMaxLd || ldb > kMaxLd) {
        if (config_id >= 0) return SGEMM_STATUS_NOT_SUPPORTED;
        return hip_status(launch_wide(p, stream));
    }
    if (config_id < 0) config_id = choose_config(form, m, n, k, columns_16b_aligned(p));
    hipError_t e = form == SGEMM_FORM_F32 ? launch_f32(config_id, p, stream, arena)
                                          : launch_bf16x3(config_id, p, stream, arena);
    // With an arena nothing is allocated: out of memory means the plan needs
    // more partials or counters than the arena holds (nothing was launched).
    if (arena != nullptr && e == hipErrorOutOfMemory) return SGEMM_STATUS_SCRATCH_TOO_SMALL;
    return hip_status(e);
}

namespace {

int run(int form, int config_id, int m, int n, int k, const float* alpha, const float* A, int lda,
        const float* B, int ldb, const float* beta, float* C, int ldc, hipStream_t stream) {
    return run_problem(form, config_id, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream,
                       nullptr);
}

}  // namespace

extern "C" {

void sgemm(int m, int n, int k, const float* alpha, float* A, int lda, float* B, int ldb,
           const float* beta, float* C, int ldc) {
    record(run(SGEMM_FORM_F32, -1, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nullptr));
}

void sgemm_basic(int m, int n, int k, const float* alpha, float* A, int lda, float* B, int ldb,
                 const float* beta, float* C, int ldc) {
    int st = validate(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    if (st == SGEMM_STATUS_SUCCESS && m > 0 && n > 0) {
        if (k > 0 && (A == nullptr || B == nullptr)) {
            st = SGEMM_STATUS_INVALID_VALUE;
        } else {
            Problem p = make_problem(m, n, k, *alpha, A, lda, B, ldb, *beta, C, ldc);
            st = hip_status(launch_basic(p, nullptr));
        }
    }
    record(st);
}

void sgemm_rm(int m, int n, int k, const float* alpha, float* A, int lda, float* B, int ldb,
              const float* beta, float* C, int ldc) {
    // Row-major C (m x n) is column-major C^T (n x m) = B^T A^T.
    record(run(SGEMM_FORM_F32, -1, n, m, k, alpha, B, ldb, A, lda, beta, C, ldc, nullptr));
}

int sgemm_get_last_error(void) {
    int e = g_last_error;
    g_last_error = SGEMM_STATUS_SUCCESS;
    return e;
}

const char* sgemm_status_string(int status) {
    switch (status) {
        case SGEMM_STATUS_SUCCESS: return "success";
        case SGEMM_STATUS_INVALID_VALUE: return "invalid value";
        case SGEMM_STATUS_NOT_SUPPORTED: return "not supported";
        case SGEMM_STATUS_HIP_ERROR: return "HIP runtime error";
        case SGEMM_STATUS_ALLOC_FAILED: return "allocation failed";
        case SGEMM_STATUS_BUDGET_EXCEEDED: return "request exceeds the HBM budget";
        case SGEMM_STATUS_NOT_READY: return "not ready";
        case SGEMM_STATUS_NOT_FOUND: return "not found";
        case SGEMM_STATUS_IO_ERROR: return "I/O error";
        case SGEMM_STATUS_WRONG_STATE: return "wrong state";
        case SGEMM_STATUS_SCRATCH_TOO_SMALL:
            return "the problem needs more scratch than the static no-allocation array holds";
        default: return "unknown status";
    }
}

int sgemm_ex(int m, int n, int k, const float* alpha, const float* A, int lda, const float* B,
             int ldb, const float* beta, float* C, int ldc, sgemm_stream_t stream) {
    return run(SGEMM_FORM_F32, -1, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
               static_cast<hipStream_t>(stream));
}

int sgemm_bf16x3(int m, int n, int k, const float* alpha, const float* A, int lda,
                 const float* B, int ldb, const float* beta, float* C, int ldc,
                 sgemm_stream_t stream) {
    return run(SGEMM_FORM_BF16X3, -1, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
               static_cast<hipStream_t>(stream));
}

int sgemm_run_config(int form, int config_id, int m, int n, int k, const float* alpha,
                     const float* A, int lda, const float* B, int ldb, const float* beta,
                     float* C, int ldc, sgemm_stream_t stream) {
    if (config_id < 0) return SGEMM_STATUS_NOT_SUPPORTED;
    return run(form, config_id, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
               static_cast<hipStream_t>(stream));
}

int sgemm_config_count(int form) { return config_count(form); }

int sgemm_launch_config(int form, int m, int n, int k, const float* A, int lda, const float* B,
                        int ldb, int* config_id) {
    if (form != SGEMM_FORM_F32 && form != SGEMM_FORM_BF16X3) return SGEMM_STATUS_NOT_SUPPORTED;
    if (config_id == nullptr || m < 0 || n < 0 || k < 0) return SGEMM_STATUS_INVALID_VALUE;
    *config_id = -1;
    // As run_problem: nothing to multiply, or leading dimensions past the
    // tiles' 32-bit offsets, launch no tile config.
    if (m == 0 || n == 0 || k == 0 || lda > kMaxLd || ldb > kMaxLd) return SGEMM_STATUS_SUCCESS;
    const Problem p = make_problem(m, n, k, 1.f, A, lda, B, ldb, 0.f, nullptr, m);
    *config_id = choose_config(form, m, n, k, columns_16b_aligned(p));
    return SGEMM_STATUS_SUCCESS;
}

int sgemm_config_info(int form, int config_id, int* bm, int* bn, int* bk, char* name,
                      int name_len) {
    if (config_id < 0 || config_id >= config_count(form)) return SGEMM_STATUS_NOT_SUPPORTED;
    const ConfigInfo& c = config_info(form, config_id);
    if (bm) *bm = c.bm;
    if (bn) *bn = c.bn;
    if (bk) *bk = c.bk;
    if (name && name_len > 0) std::snprintf(name, (size_t)name_len, "%s", c.name);
    return SGEMM_STATUS_SUCCESS;
}

const char* sgemm_hip_error_string(void) {
    return g_last_hip_error == hipSuccess ? "no HIP error" : hipGetErrorString(g_last_hip_error);
}

const char* sgemm_version(void) { return "sgemm_amd 0.1 (gfx950)"; }

}  // extern "C"
