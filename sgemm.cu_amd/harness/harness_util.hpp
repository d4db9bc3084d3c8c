// harness_util.hpp -- host utilities of the C++ test / benchmark executables:
// --key=value flags, seeded inputs, HIP error checks, cache flush.
//
// Mirrors the reference's helpers (common/helper_string.h:17-65 flag parsing,
// common/helper_matrix.h:20-86 init / compare, common/helper_cuda.h:55-83
// error check and L2 flush) for HIP, with two deliberate changes: inputs are
// seeded (SplitMix64) instead of time-seeded rand(), and the flush covers the
// 256 MB Infinity Cache as well as L2.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace harness {

#define HIP_CHECK(expr)                                                                   \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            std::fprintf(stderr, "HIP error at %s:%d: %s (%s)\n", __FILE__, __LINE__,     \
                         hipGetErrorString(e_), #expr);                                   \
            std::exit(EXIT_FAILURE);                                                      \
        }                                                                                 \
    } while (0)

// Case-insensitive --key=value lookup; `def` when absent or empty
// (semantics of get_cmd_line_arg_int/_string, common/helper_string.h:17-65).
inline std::string flag(int argc, char** argv, const char* key, const std::string& def) {
    const size_t kl = std::strlen(key);
    for (int i = 1; i < argc; ++i) {
        const char* a = argv[i];
        if (a[0] != '-' || a[1] != '-') continue;
        if (strncasecmp(a + 2, key, kl) == 0 && a[2 + kl] == '=') {
            const char* v = a + 3 + kl;
            return *v ? std::string(v) : def;
        }
    }
    return def;
}

inline int flag_int(int argc, char** argv, const char* key, int def) {
    std::string v = flag(argc, argv, key, "");
    return v.empty() ? def : std::atoi(v.c_str());
}

inline bool has_flag(int argc, char** argv, const char* key) {
    for (int i = 1; i < argc; ++i)
        if (argv[i][0] == '-' && argv[i][1] == '-' && strcasecmp(argv[i] + 2, key) == 0)
            return true;
    return false;
}

// SplitMix64 (counter form): output i of the stream seeded with `seed`.
inline uint64_t splitmix64_at(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Uniform [0, 1) with 24 random bits: the reference's input range
// (init_random, common/helper_matrix.h:20-23), reproducibly seeded.
inline void init_unit(std::vector<float>& x, uint64_t seed) {
    for (size_t i = 0; i < x.size(); ++i)
        x[i] = (float)((double)(splitmix64_at(seed, i) >> 40) * (1.0 / 16777216.0));
}

struct CompareResult {
    bool equal;
    long long n_false, n_nan, n_inf;
};

// Relative error |ref - x| / |ref| > eps fails; NaN and Inf in x always fail
// (compare_mats, common/helper_matrix.h:51-86).
inline CompareResult compare(const float* ref, const float* x, size_t n, double eps) {
    CompareResult r{true, 0, 0, 0};
    for (size_t i = 0; i < n; ++i) {
        if (std::isnan(x[i])) {
            ++r.n_nan;
            ++r.n_false;
            continue;
        }
        if (std::isinf(x[i])) {
            ++r.n_inf;
            ++r.n_false;
            continue;
        }
        if (std::fabs(((double)ref[i] - (double)x[i]) / (double)ref[i]) > eps) ++r.n_false;
    }
    r.equal = r.n_false == 0;
    return r;
}

// Evicts A, B, C from L2 and the Infinity Cache between timed calls by writing
// a buffer larger than both (512 MiB), outside the timed region
// (the reference flushes L2 only: l2flush, common/helper_cuda.h:70-83).
class CacheFlusher {
  public:
    CacheFlusher() { HIP_CHECK(hipMalloc(&buf_, kBytes)); }
    ~CacheFlusher() { (void)hipFree(buf_); }
    void flush(hipStream_t s = nullptr) {
        HIP_CHECK(hipMemsetAsync(buf_, ++val_ & 0xff, kBytes, s));
    }

  private:
    static constexpr size_t kBytes = 512ull << 20;
    void* buf_ = nullptr;
    int val_ = 0;
};

}  // namespace harness
