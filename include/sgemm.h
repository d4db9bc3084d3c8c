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
    SGEMM_STATUS_WRONG_STATE = 9,     /* e.g. replay of a plan that was never captured */
    SGEMM_STATUS_SCRATCH_TOO_SMALL = 10 /* the problem's split plan needs more than a fixed
                                           scratch array holds; nothing was launched */
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
 * launch nothing and are reported by sgemm_get_last_error().  lda or ldb
 * above 2^21 (past the tiled kernels' 32-bit buffer offsets) run a simple
 * FP32 kernel with 64-bit addressing -- correct, not fast; sgemm_run_config
 * rejects them (SGEMM_STATUS_NOT_SUPPORTED).
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
/* Text of the last HIP runtime error behind a SGEMM_STATUS_HIP_ERROR on this
 * host thread. */
const char* sgemm_hip_error_string(void);

/* ------------------------------------------------------------------------ */
/* Status-returning, stream-aware forms (no reference counterpart; the       */
/* reference's only entry point is sgemm() above).                           */
/* ------------------------------------------------------------------------ */

/* As sgemm(), on `stream`, returning a status.  The config is chosen by the
 * autotune table for (m, n, k, SGEMM_FORM_F32).  Split-K configs (chosen for
 * grids that would hold few workgroups per CU) and the tail-split config
 * (for grids whose last round of resident workgroups is partial) use a device
 * workspace owned by the library, one per stream, allocated on the first such
 * call and kept until sgemm_workspace_release(stream) or the process ends; a
 * call on a stream under graph capture never allocates and runs unsplit if
 * the workspace is too small.  Results of a split config are deterministic
 * (partials are summed in split order).  Launches that share a stream's
 * workspace must not overlap: a graph that captured such calls on `stream`
 * uses that stream's workspace whenever it is replayed, so its replays must
 * be ordered with each other and with eager calls on `stream`. */
int sgemm_ex(int m, int n, int k, const float* alpha, const float* A, int lda, const float* B,
             int ldb, const float* beta, float* C, int ldc, sgemm_stream_t stream);

/* Split-precision form: each FP32 input x is split on the GPU into
 * hi = bf16_rne(x), lo = bf16_rne(x - hi); A*B is accumulated in FP32 as
 * hi*hi + hi*lo + lo*hi on BF16 MFMA.  Same operand conventions as sgemm_ex.
 * Besides split-K and tail-split configs this form has a stream-K config
 * (the tiles of a partial round spread k-tile by k-tile over the CUs, cut
 * tiles summed from their pieces in k order, so also deterministic), with the
 * same per-stream workspace and capture behaviour; its partials take up to
 * 512 MiB.  For large problems it splits A and B once, in a pass of their own,
 * into packed hi/lo planes held in the stream's workspace (about the size of
 * A and B, up to 16 GiB; config 12); where that workspace cannot be had (a
 * stream under capture whose workspace would have to grow, an allocation
 * failure, larger problems) the same tile runs from A and B directly (config
 * 2), with the same result bits. */
int sgemm_bf16x3(int m, int n, int k, const float* alpha, const float* A, int lda,
                 const float* B, int ldb, const float* beta, float* C, int ldc,
                 sgemm_stream_t stream);

/* No-allocation form of sgemm_ex (form SGEMM_FORM_F32) and sgemm_bf16x3
 * (form SGEMM_FORM_BF16X3).  No counterpart in the reference (its sgemm keeps
 * no device state); for callers that cannot allocate device memory: code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  Same arguments, config choice, kernels and result bits as those
 * entry points (except that a packed BF16x3 pick, config 12, runs config 2's
 * tile from A and B, which gives the same bits), but the split-K, tail-split
 * and stream-K partials and arrival
 * counters come from one static device array per GPU, held until the process
 * ends (SGEMM_NOALLOC_SCRATCH_PARTIAL_BYTES of partials and
 * SGEMM_NOALLOC_SCRATCH_COUNTERS counters), instead of a per-stream
 * workspace: nothing is allocated, also under capture (where sgemm_ex would
 * run such a problem unsplit).  A problem whose split plan needs more than
 * the array holds returns SGEMM_STATUS_SCRATCH_TOO_SMALL and launches
 * nothing.  Every call on one GPU shares that GPU's array: calls that overlap
 * in time on one GPU -- on other streams, or replays of graphs that captured
 * a call -- race on it and must be ordered by the caller. */
#define SGEMM_NOALLOC_SCRATCH_PARTIAL_BYTES (64u << 20)
#define SGEMM_NOALLOC_SCRATCH_COUNTERS 16384
int sgemm_ex_noalloc(int form, int m, int n, int k, const float* alpha, const float* A, int lda,
                     const float* B, int ldb, const float* beta, float* C, int ldc,
                     sgemm_stream_t stream);

/* Frees the split-K / stream-K workspace and the packed BF16x3 planes the
 * library holds for `stream` (waiting for the stream first).  Call it before
 * destroying a stream that ran split or packed configs, so the handle cannot
 * be reused while its workspace is still in use; graphs that captured such
 * launches on `stream` must not be replayed afterwards.  A stream without a workspace is a no-op.  No
 * reference counterpart (the reference holds no device state). */
int sgemm_workspace_release(sgemm_stream_t stream);

/* Explicit form and kernel config (config ids from sgemm_config_count /
 * sgemm_config_info; splitk configs as described at sgemm_ex).  Used by the
 * autotuner and by tests. */
int sgemm_run_config(int form, int config_id, int m, int n, int k, const float* alpha,
                     const float* A, int lda, const float* B, int ldb, const float* beta,
                     float* C, int ldc, sgemm_stream_t stream);
int sgemm_config_count(int form);
/* Config id that sgemm_ex (form SGEMM_FORM_F32) or sgemm_bf16x3 would launch
 * for this problem with these operands (the table or heuristic choice, which
 * depends on the shape and, for FP32, on whether A, B, lda and ldb give
 * 16-byte aligned columns); -1 where no tile config runs (k == 0, an empty
 * problem, or a leading dimension past 2^21, which take the scale-only or
 * the naive kernel).  Reads no memory; no reference counterpart. */
int sgemm_launch_config(int form, int m, int n, int k, const float* A, int lda, const float* B,
                        int ldb, int* config_id);
/* Writes "BMxBNxBK/WMxWN" style text; returns the tile sizes. */
int sgemm_config_info(int form, int config_id, int* bm, int* bn, int* bk, char* name,
                      int name_len);

/* ------------------------------------------------------------------------ */
/* Grouped GEMM: many independent problems in one persistent launch (no     */
/* reference counterpart).                                                   */
/* ------------------------------------------------------------------------ */

/* One problem of a group: column-major, device operands, scalars by value.
 * Same argument rules as sgemm_ex. */
typedef struct {
    int m, n, k;
    float alpha, beta;
    const float* A;
    int lda;
    const float* B;
    int ldb;
    float* C;
    int ldc;
} sgemm_problem;

typedef struct sgemm_grouped_plan_s* sgemm_grouped_plan;

/* A plan holds device descriptors for up to `capacity` problems of one form. */
int sgemm_grouped_plan_create(sgemm_grouped_plan* plan, int capacity, int form);
/* Validate, pick each problem's tile config from the autotune table, order the
 * tiles (most expensive first) and upload the descriptors on `stream`.  The
 * problems array is copied; it may be reused when this returns. */
int sgemm_grouped_plan_set(sgemm_grouped_plan plan, int count, const sgemm_problem* problems,
                           sgemm_stream_t stream);
/* As sgemm_grouped_plan_set with an explicit config id per problem. */
int sgemm_grouped_plan_set_configs(sgemm_grouped_plan plan, int count,
                                   const sgemm_problem* problems, const int* config_ids,
                                   sgemm_stream_t stream);
/* One persistent kernel launch computing every problem of the plan.  Its
 * workgroups take tiles from a queue in the plan's device memory, which the
 * launch leaves empty again, so runs and replays of one plan must be ordered
 * (one stream, or events between streams): two overlapping launches of the
 * same plan would share its queue.  Different plans may run concurrently. */
int sgemm_grouped_plan_run(sgemm_grouped_plan plan, sgemm_stream_t stream);
/* Capture the launch into a hipGraph (stream capture on `stream`) ... */
int sgemm_grouped_plan_capture(sgemm_grouped_plan plan, sgemm_stream_t stream);
/* ... and launch that graph.  The graph reads the plan's current descriptors,
 * so a later sgemm_grouped_plan_set (same plan) needs no new capture. */
int sgemm_grouped_plan_replay(sgemm_grouped_plan plan, sgemm_stream_t stream);
int sgemm_grouped_plan_info(sgemm_grouped_plan plan, int* count, int* total_tiles, int* grid,
                            double* flops);
/* Waits for the device, then frees the plan. */
int sgemm_grouped_plan_destroy(sgemm_grouped_plan plan);
/* One-shot: plan, run on `stream`, wait for it, free. */
int sgemm_grouped(int count, const sgemm_problem* problems, int form, sgemm_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Autotuner (no reference counterpart): one tile config per (form, grouped, */
/* size bucket of m, n, k), bucket(x) = ceil(log2 x) capped to [0, 15].      */
/* Selection is a pure function of the shape, the loaded table and, for the */
/* FP32 heuristic, whether every column of A and B is 16-byte aligned (A, B */
/* 16-byte aligned, lda and ldb multiples of 4: configs 5, 9 and 10 then    */
/* load their operands straight into LDS and hold four workgroups per CU).  */
/* A table named by the environment variable SGEMM_AMD_TUNE_TABLE loads on  */
/* first use.                                                               */
/* ------------------------------------------------------------------------ */

/* Time every candidate config on an (m, n, k) problem (`iters` launches each,
 * after one warm-up) and record the fastest for that shape's bucket.
 * Allocates its own operands; not for hot paths. */
int sgemm_autotune_bucket(int form, int grouped, int m, int n, int k, int iters,
                          sgemm_stream_t stream, int* best_config, float* best_tflops);
/* Config the library will use for this shape (with 16-byte aligned operand
 * columns); *from_table = 1 if a table entry (not the built-in heuristic)
 * decided. */
int sgemm_autotune_lookup(int form, int grouped, int m, int n, int k, int* config_id,
                          int* from_table);
/* Set one bucket's config by hand. */
int sgemm_autotune_set(int form, int grouped, int m, int n, int k, int config_id);
int sgemm_autotune_clear(void);
int sgemm_autotune_size(void);
/* Persist / restore the table (JSON, one entry per line). */
int sgemm_autotune_save(const char* path);
int sgemm_autotune_load(const char* path);

/* ------------------------------------------------------------------------ */
/* Serving loop: continuous batching of arriving GEMM requests into grouped  */
/* launches, admitted against a fixed HBM budget (no reference counterpart). */
/* Small requests share grouped launches; large ones are cut into column    */
/* panels that are interleaved with those launches (see chunk_gflop).       */
/* ------------------------------------------------------------------------ */

enum { SGEMM_EXECUTOR_GPU = 0, SGEMM_EXECUTOR_MOCK = 1 };
enum { SGEMM_OPERANDS_DEVICE = 0, SGEMM_OPERANDS_HOST = 1 };

typedef struct {
    size_t hbm_budget_bytes; /* admission budget; with host operands also the pool size */
    int max_batch;           /* requests per grouped launch */
    int num_slots;           /* grouped launches in flight at once (one stream each), 1..16 */
    int form;                /* SGEMM_FORM_F32 or SGEMM_FORM_BF16X3 */
    int executor;            /* SGEMM_EXECUTOR_GPU, or _MOCK (no device: a batch completes
                                after flops / mock_tflops seconds) */
    double mock_tflops;
    double mock_launch_us;
    double chunk_gflop;      /* device-operand requests above this many GFLOP (2mnk/1e9) are
                                cut into column panels of about this size, each one launch
                                interleaved with the small-request batches (the config the
                                autotune table picks for the panel's shape, except that a
                                tail-split or stream-K pick runs the same tile whole, since
                                other slots share the GPU, and a packed BF16x3 pick runs
                                config 2, whose bits it shares); 0 = 100 GFLOP, < 0 = never cut
                                (every request goes into grouped batches) */
    int mock_fail_launch;    /* SGEMM_EXECUTOR_MOCK only: the launch with this 1-based index
                                (grouped batch or panel) fails, to test failure handling;
                                0 = none */
} sgemm_server_config;

/* One request.  operands == SGEMM_OPERANDS_DEVICE: A, B, C are device memory
 * the caller keeps alive until the request has finished (poll/wait return a
 * status other than SGEMM_STATUS_NOT_READY, success or error: a failed
 * request's status is published only after all of its launches have left
 * the GPU).  SGEMM_OPERANDS_HOST: host memory; the server copies A, B (and C
 * if beta != 0) into a device pool of hbm_budget_bytes, allocated by the
 * first such request, on admission and C back on completion. */
typedef struct {
    int m, n, k;
    float alpha, beta;
    const float* A;
    int lda;
    const float* B;
    int ldb;
    float* C;
    int ldc;
    int operands;
} sgemm_request;

typedef struct {
    long long submitted, completed, rejected, failed;
    long long batches;
    long long measured;   /* completed requests since the last reset; mean, max
                             and tflops cover all of them, p50 and p99 the
                             most recent 65536 */
    double mean_batch;    /* requests per grouped launch */
    double p50_ms, p99_ms, mean_ms, max_ms; /* completion - submission */
    double mean_queue_ms;                    /* admission - submission */
    double window_s;      /* first submission .. last completion */
    double tflops;        /* sum of 2mnk over that window */
    size_t peak_bytes;    /* most budget bytes in use at once */
    size_t budget_bytes;
    long long panel_launches; /* single-problem launches of large-request panels */
} sgemm_server_stats;

typedef struct sgemm_server_s* sgemm_server;

int sgemm_server_create(sgemm_server* srv, const sgemm_server_config* cfg);
/* Queue a request; *id names it.  A request whose footprint
 * 4*(mk + kn + mn) bytes exceeds the budget is rejected with
 * SGEMM_STATUS_BUDGET_EXCEEDED. */
int sgemm_server_submit(sgemm_server srv, const sgemm_request* req, uint64_t* id);
/* SGEMM_STATUS_NOT_READY while queued or in flight, then its final status.
 * The last 65536 finished requests stay queryable (poll, wait, latency);
 * older ids return SGEMM_STATUS_NOT_FOUND. */
int sgemm_server_poll(sgemm_server srv, uint64_t id);
int sgemm_server_wait(sgemm_server srv, uint64_t id, double timeout_s);
/* Wait until every submitted request has finished. */
int sgemm_server_drain(sgemm_server srv, double timeout_s);
int sgemm_server_latency(sgemm_server srv, uint64_t id, double* total_ms, double* queue_ms);
int sgemm_server_get_stats(sgemm_server srv, sgemm_server_stats* out);
/* Latency statistics then cover requests submitted after this call. */
int sgemm_server_reset_stats(sgemm_server srv);
/* Finishes in-flight launches, then frees everything. */
int sgemm_server_destroy(sgemm_server srv);

/* Library version string. */
const char* sgemm_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SGEMM_AMD_SGEMM_H */
