// noalloc.hip -- sgemm_ex_noalloc: the single-problem entry points without any
// device allocation, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).  No
// reference counterpart: the reference's sgemm (src/sgemm.cuh:19-48) keeps
// no device state at all.
//
// The split-K, tail-split and stream-K configs need partials and per-tile
// arrival counters.  sgemm_ex / sgemm_bf16x3 take them from a per-stream
// workspace that grows on demand; this entry point takes them from one
// file-scope device array, `scratch`, of a fixed size, that every GPU the
// library runs on holds until the process ends.  The config choice, kernels
// and summation order are those of sgemm_ex / sgemm_bf16x3, so the results
// are the same bits; a problem whose plan needs more than the array holds is
// rejected with SGEMM_STATUS_SCRATCH_TOO_SMALL before anything is launched.
//
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address is looked up once per device (hipGetSymbolAddress, with the
// device current) and only that address is passed to kernels.  The counters
// must be zero before a split launch; every split launch leaves them zero,
// and the first call on each device zeroes them with a fill kernel on the
// caller's stream (a kernel, not a memset, so a captured graph holds no
// memset node).  All calls on one GPU share the array: calls that overlap in
// time (on other streams, or replays of graphs that captured a call) race
// on it and must be ordered by the caller.
#include <hip/hip_runtime.h>

#include <mutex>

#include "sgemm.h"
#include "kernels.hpp"

namespace sgemm_amd {
namespace {

// Arrival counters first (one int per split tile or stream-K workgroup),
// then the partial sums, 256-byte aligned.
constexpr size_t kScratchCounters = SGEMM_NOALLOC_SCRATCH_COUNTERS;
constexpr size_t kScratchCounterBytes = kScratchCounters * sizeof(int);
constexpr size_t kScratchPartialBytes = SGEMM_NOALLOC_SCRATCH_PARTIAL_BYTES;
constexpr size_t kScratchBytes = kScratchCounterBytes + kScratchPartialBytes;
static_assert(kScratchCounterBytes % 256 == 0, "partials stay 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

__global__ __launch_bounds__(256) void k_zero_counters(int* __restrict__ c, int n) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) c[i] = 0;
}

constexpr int kMaxDevices = 64;

struct DeviceScratch {
    unsigned char* ptr = nullptr;  // device address of `scratch` on this device
    bool zeroed = false;           // counters zeroed by an eager (uncaptured) call
};

std::mutex g_mu;
DeviceScratch g_dev[kMaxDevices];

// The current device's scratch as an arena; zeroes the counters on `stream`
// on the first call on the device (again on later calls while every earlier
// one was being captured: a graph's fill runs only when it is replayed).
hipError_t device_arena(hipStream_t stream, Arena* out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    DeviceScratch& d = g_dev[dev];
    if (d.ptr == nullptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        d.ptr = static_cast<unsigned char*>(p);
    }
    if (!d.zeroed) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(stream, &cs) != hipSuccess) {
            (void)hipGetLastError();
            cs = hipStreamCaptureStatusNone;
        }
        (void)hipGetLastError();
        hipLaunchKernelGGL(k_zero_counters, dim3(16), dim3(256), 0, stream,
                           reinterpret_cast<int*>(d.ptr), (int)kScratchCounters);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        if (cs == hipStreamCaptureStatusNone) d.zeroed = true;
    }
    out->counters = reinterpret_cast<int*>(d.ptr);
    out->counter_count = (int)kScratchCounters;
    out->partials = reinterpret_cast<float*>(d.ptr + kScratchCounterBytes);
    out->partial_bytes = kScratchPartialBytes;
    return hipSuccess;
}

}  // namespace
}  // namespace sgemm_amd

extern "C" int sgemm_ex_noalloc(int form, int m, int n, int k, const float* alpha,
                                const float* A, int lda, const float* B, int ldb,
                                const float* beta, float* C, int ldc, sgemm_stream_t stream) {
    using namespace sgemm_amd;
    const int st = validate_args(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    if (st != SGEMM_STATUS_SUCCESS) return st;
    if (form != SGEMM_FORM_F32 && form != SGEMM_FORM_BF16X3) return SGEMM_STATUS_NOT_SUPPORTED;
    if (m == 0 || n == 0) return SGEMM_STATUS_SUCCESS;
    hipStream_t s = static_cast<hipStream_t>(stream);
    Arena arena{};
    const hipError_t e = device_arena(s, &arena);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return SGEMM_STATUS_HIP_ERROR;
    }
    return run_problem(form, -1, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, &arena);
}
