// serve_stress.cpp -- the serving C ABI under concurrent use, on the mock
// executor (no GPU), for the thread / address / undefined-behaviour
// sanitizer builds of tests/sanitize/Makefile.
//
// Four threads submit mixed requests (small ones for grouped batches, large
// ones cut into panels, host-operand ones carved out of the pool), two poll,
// wait on and query the latency of ids already handed out, one reads and
// resets the statistics, while the scheduler thread admits, launches (one
// injected launch failure included) and retires.  A second server is
// destroyed with requests still queued and in flight.  Exit status 0 when
// every request is accounted for (completed + failed == submitted) and the
// statuses are consistent; the sanitizers report anything else themselves.
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <random>
#include <thread>
#include <vector>

#include "sgemm.h"

namespace {

int fail(const char* what, long long a, long long b) {
    std::fprintf(stderr, "serve_stress: %s (%lld vs %lld)\n", what, a, b);
    return 1;
}

sgemm_server_config mock_config(size_t budget, int fail_launch) {
    sgemm_server_config c{};
    c.hbm_budget_bytes = budget;
    c.max_batch = 8;
    c.num_slots = 3;
    c.form = SGEMM_FORM_BF16X3;
    c.executor = SGEMM_EXECUTOR_MOCK;
    c.mock_tflops = 0.2;
    c.mock_launch_us = 5;
    c.chunk_gflop = 0.01;  // requests above 10 MFLOP run as column panels
    c.mock_fail_launch = fail_launch;
    return c;
}

// A dummy device pointer: the mock executor never dereferences operands.
float* const kDummy = reinterpret_cast<float*>(uintptr_t{4096});

int submit_random(sgemm_server srv, std::mt19937& rng, uint64_t* id) {
    std::uniform_int_distribution<int> pick(0, 99), small(8, 96), large(160, 320);
    const int r = pick(rng);
    sgemm_request q{};
    const bool big = r < 10;
    q.m = big ? large(rng) : small(rng);
    q.n = big ? 1024 : small(rng);
    q.k = big ? large(rng) : small(rng);
    q.alpha = 1.f;
    q.beta = (r % 3 == 0) ? 0.5f : 0.f;
    q.A = kDummy;
    q.lda = q.m;
    q.B = kDummy;
    q.ldb = q.k;
    q.C = kDummy;
    q.ldc = q.m;
    q.operands = (!big && r % 5 == 0) ? SGEMM_OPERANDS_HOST : SGEMM_OPERANDS_DEVICE;
    return sgemm_server_submit(srv, &q, id);
}

}  // namespace

int main() {
    sgemm_server srv = nullptr;
    sgemm_server_config cfg = mock_config(64u << 20, 37);
    if (sgemm_server_create(&srv, &cfg) != SGEMM_STATUS_SUCCESS) return fail("create", 0, 0);

    constexpr int kSubmitters = 4, kPerThread = 1500;
    std::atomic<uint64_t> max_id{0};
    std::atomic<bool> done{false};
    std::atomic<long long> submitted{0};
    std::vector<std::thread> th;
    for (int t = 0; t < kSubmitters; ++t)
        th.emplace_back([&, t] {
            std::mt19937 rng(1000 + t);
            for (int i = 0; i < kPerThread; ++i) {
                uint64_t id = 0;
                if (submit_random(srv, rng, &id) == SGEMM_STATUS_SUCCESS) {
                    ++submitted;
                    uint64_t m = max_id.load();
                    while (id > m && !max_id.compare_exchange_weak(m, id)) {
                    }
                }
                if (i % 64 == 0) std::this_thread::yield();
            }
        });
    std::atomic<long long> bad{0};
    std::vector<std::thread> side;
    for (int t = 0; t < 2; ++t)
        side.emplace_back([&, t] {
            std::mt19937 rng(2000 + t);
            while (!done.load()) {
                const uint64_t m = max_id.load();
                if (m == 0) continue;
                const uint64_t id = 1 + rng() % m;
                const int st = (t == 0) ? sgemm_server_poll(srv, id)
                                        : sgemm_server_wait(srv, id, 0.001);
                if (st != SGEMM_STATUS_SUCCESS && st != SGEMM_STATUS_NOT_READY &&
                    st != SGEMM_STATUS_HIP_ERROR)
                    ++bad;
                double tot = 0, q = 0;
                const int ls = sgemm_server_latency(srv, id, &tot, &q);
                if (ls == SGEMM_STATUS_SUCCESS && !(tot >= q && q >= 0)) ++bad;
            }
        });
    side.emplace_back([&] {
        int n = 0;
        while (!done.load()) {
            sgemm_server_stats s;
            sgemm_server_get_stats(srv, &s);
            if (s.completed + s.failed > s.submitted) ++bad;
            if (++n % 200 == 0) sgemm_server_reset_stats(srv);
            std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
    });
    for (auto& x : th) x.join();
    const int drained = sgemm_server_drain(srv, 120.0);
    done = true;
    for (auto& x : side) x.join();
    if (drained == SGEMM_STATUS_NOT_READY) return fail("drain timed out", 0, 0);
    sgemm_server_stats s;
    sgemm_server_get_stats(srv, &s);
    if (s.submitted != submitted.load()) return fail("submitted", s.submitted, submitted.load());
    if (s.completed + s.failed != s.submitted)
        return fail("completed + failed != submitted", s.completed + s.failed, s.submitted);
    if (s.failed < 1) return fail("the injected launch failure was not reported", s.failed, 1);
    if (bad.load() != 0) return fail("inconsistent status or latency", bad.load(), 0);
    if (sgemm_server_destroy(srv) != SGEMM_STATUS_SUCCESS) return fail("destroy", 0, 0);

    // Destroy with work queued and in flight.
    cfg = mock_config(8u << 20, 0);
    cfg.mock_tflops = 0.001;
    if (sgemm_server_create(&srv, &cfg) != SGEMM_STATUS_SUCCESS) return fail("create 2", 0, 0);
    std::mt19937 rng(7);
    for (int i = 0; i < 200; ++i) {
        uint64_t id = 0;
        (void)submit_random(srv, rng, &id);
    }
    if (sgemm_server_destroy(srv) != SGEMM_STATUS_SUCCESS) return fail("destroy 2", 0, 0);
    std::printf("serve_stress: %lld requests, %lld failed by injection, ok\n", s.submitted,
                s.failed);
    return 0;
}
