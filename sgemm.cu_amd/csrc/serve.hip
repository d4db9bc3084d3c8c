// serve.hip -- continuous-batching GEMM server (no reference counterpart; the
// reference has no serving layer).
//
// Requests arrive through sgemm_server_submit.  A scheduler thread admits
// them against a fixed HBM budget and launches them on "slots"; every slot has
// its own stream, so up to num_slots launches overlap on the device.  A
// request that arrives while others are on the GPU is admitted as soon as a
// slot and enough of the budget are free -- it does not wait for the running
// launches.
//
// Two lanes, each strict FIFO:
//   * small requests (2mnk <= chunk threshold, and every host-operand request)
//     are packed, up to max_batch at a time, into one grouped launch
//     (grouped.hip);
//   * large device-operand requests are cut into column panels
//     C[:, j0:j1] = alpha*A*B[:, j0:j1] + beta*C[:, j0:j1] of about the
//     threshold's flops each, and every panel is one single-problem launch with
//     the best config for its shape.  Panels of one large request go out in
//     order; the next large request starts when the previous one has launched
//     all of its panels.
// A free slot alternates between the lanes (a small batch, then a panel, ...),
// falling back to whichever lane has work.  So a small request waits for at
// most about one panel of large work instead of a whole large request, and a
// large request still advances by a panel every other launch: neither lane
// starves.  A small request may overtake a large one waiting for memory only
// if the budget still has room for that large one afterwards.
// Footprint = 4*(m*k + k*n + m*n) bytes (the request's operands and result),
// reserved from admission to completion.
//   * SGEMM_OPERANDS_HOST: the server owns the device copies.  It carves them
//     out of one pool of the budget's size, allocated by the first such
//     request (first-fit, coalescing free list), copies A, B (and C when
//     beta != 0) in on the slot's stream, and C back after the launch.
//   * SGEMM_OPERANDS_DEVICE: the operands already live on the device; the
//     footprint is reserved against the budget while the request is in flight.
//
// Which lane, which panels and which config a request gets depend on its own
// shape only, so its bits do not depend on which other requests share the GPU.
//
// A request's status is published once all of its work has left the GPU:
// when one panel of a large request fails, its panels not yet launched are
// dropped and the error is reported after the panels still in flight have
// retired, so a caller that frees the operands on the error is safe.
// Finished requests stay queryable (poll, wait, latency) for the last 65536
// of them; the statistics are kept as running figures since the last reset.
//
// The executor is replaceable: SGEMM_EXECUTOR_MOCK completes a launch after
// flops / mock_tflops seconds without touching a GPU, so admission, batching,
// failure handling and the statistics are testable on any host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "sgemm.h"
#include "grouped.hpp"

namespace sgemm_amd {
namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// First-fit allocator over [0, capacity) with coalescing; offsets only, so the
// mock executor can use it without device memory.
class RangeAllocator {
  public:
    explicit RangeAllocator(size_t capacity) { free_[0] = capacity; }
    bool alloc(size_t bytes, size_t* off) {
        bytes = align_up(std::max<size_t>(bytes, 1), 256);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes;
                const size_t next = it->first + bytes;
                free_.erase(it);
                if (rest) free_[next] = rest;
                used_[*off] = bytes;
                return true;
            }
        }
        return false;
    }
    void release(size_t off) {
        auto u = used_.find(off);
        if (u == used_.end()) return;
        size_t start = off, len = u->second;
        used_.erase(u);
        auto next = free_.lower_bound(start);
        if (next != free_.end() && start + len == next->first) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) {
                prev->second += len;
                return;
            }
        }
        free_[start] = len;
    }

  private:
    std::map<size_t, size_t> free_;  // offset -> length
    std::map<size_t, size_t> used_;
};

struct Req {
    uint64_t id = 0;
    sgemm_request r{};
    size_t footprint = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int status = SGEMM_STATUS_NOT_READY;  // published only once the request is finished
    int error = SGEMM_STATUS_SUCCESS;     // first failure seen while in flight
    bool pooled = false;
    size_t off[3] = {0, 0, 0};  // pool offsets of A, B, C (host operands)
    // Large lane: column panels of panel_n columns (the last may be narrower).
    bool large = false;
    int panels = 0, panel_n = 0, next_panel = 0, panels_done = 0;
};

// A slot is the scheduler thread's own: only that thread reads or writes it
// (the destructor runs after the thread has been joined).
struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    GroupedPlan* plan = nullptr;
    std::vector<Req*> batch;   // small lane: the grouped launch's requests
    Req* chunk = nullptr;      // large lane: the request one panel belongs to
    int chunk_panel = -1;
    bool busy = false;
    bool finished = false;     // the launch has completed (seen by poll_slots)
    bool failed = false;       // ... with an error
    int launch_status = SGEMM_STATUS_SUCCESS;
    double mock_done_at = 0;
};

// Requests above this many flops go to the large lane (when chunking is on).
constexpr double kDefaultChunkGflop = 100.0;
// Finished requests whose status and latency stay queryable.
constexpr size_t kHistory = 1 << 16;
// Latencies kept for the percentiles: the most recent this many.
constexpr size_t kLatencyWindow = 1 << 16;

}  // namespace

class Server {
  public:
    explicit Server(const sgemm_server_config& c) : cfg_(c), alloc_(c.hbm_budget_bytes) {
        chunk_flops_ = c.chunk_gflop < 0 ? 0.0
                       : (c.chunk_gflop > 0 ? c.chunk_gflop : kDefaultChunkGflop) * 1e9;
    }

    int start() {
        t0_ = Clock::now();
        const bool gpu = cfg_.executor == SGEMM_EXECUTOR_GPU;
        slots_.resize(cfg_.num_slots);
        for (Slot& s : slots_) {
            if (!gpu) continue;
            if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess)
                return SGEMM_STATUS_HIP_ERROR;
            int st = GroupedPlan::create(&s.plan, cfg_.max_batch, cfg_.form);
            if (st != SGEMM_STATUS_SUCCESS) return st;
        }
        worker_ = std::thread([this] { loop(); });
        return SGEMM_STATUS_SUCCESS;
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> lock(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        if (worker_.joinable()) worker_.join();
        for (Slot& s : slots_) {
            if (s.stream) {
                (void)hipStreamSynchronize(s.stream);
                // Single-problem panels may have grown a split-K workspace
                // for this stream; free it before the handle can be reused.
                (void)sgemm_workspace_release(s.stream);
            }
            if (s.plan) {
                s.plan->release();
                delete s.plan;
            }
            if (s.done) (void)hipEventDestroy(s.done);
            if (s.stream) (void)hipStreamDestroy(s.stream);
        }
        if (pool_) (void)hipFree(pool_);
    }

    int submit(const sgemm_request& r, uint64_t* id) {
        sgemm_problem q{r.m, r.n, r.k, r.alpha, r.beta, r.A, r.lda, r.B, r.ldb, r.C, r.ldc};
        int st = validate_problem(q);
        if (st != SGEMM_STATUS_SUCCESS) return st;
        if (r.operands != SGEMM_OPERANDS_DEVICE && r.operands != SGEMM_OPERANDS_HOST)
            return SGEMM_STATUS_INVALID_VALUE;
        auto q_ = std::make_unique<Req>();
        q_->r = r;
        q_->footprint = footprint(r);
        if (q_->footprint > cfg_.hbm_budget_bytes) {
            std::lock_guard<std::mutex> lock(mu_);
            ++rejected_;
            return SGEMM_STATUS_BUDGET_EXCEEDED;
        }
        // The device pool for host operands is allocated by the first request
        // that needs it: a server fed only device operands holds none.
        if (r.operands == SGEMM_OPERANDS_HOST) {
            st = ensure_pool();
            if (st != SGEMM_STATUS_SUCCESS) return st;
        }
        // Large lane: device operands above the chunk threshold, cut into
        // column panels of about chunk_flops_ each (multiples of 256 columns,
        // the widest tile, unless n itself is smaller).
        const double flops = 2.0 * r.m * r.n * (double)r.k;
        if (chunk_flops_ > 0 && r.operands == SGEMM_OPERANDS_DEVICE && flops > chunk_flops_) {
            long long w = (long long)(chunk_flops_ / (2.0 * r.m * (double)r.k));
            w = std::max(256LL, w / 256 * 256);
            q_->panel_n = (int)std::min<long long>(w, r.n);
            q_->panels = (r.n + q_->panel_n - 1) / q_->panel_n;
            q_->large = q_->panels > 1;
        }
        std::lock_guard<std::mutex> lock(mu_);
        q_->id = ++next_id_;
        q_->t_submit = seconds_since(t0_);
        if (id) *id = q_->id;
        (q_->large ? large_q_ : pending_).push_back(q_.get());
        all_[q_->id] = std::move(q_);
        ++submitted_;
        cv_.notify_all();
        return SGEMM_STATUS_SUCCESS;
    }

    int poll(uint64_t id) {
        std::lock_guard<std::mutex> lock(mu_);
        auto it = all_.find(id);
        return it == all_.end() ? SGEMM_STATUS_NOT_FOUND : it->second->status;
    }

    int wait(uint64_t id, double timeout_s) {
        std::unique_lock<std::mutex> lock(mu_);
        const auto deadline = Clock::now() + std::chrono::duration<double>(timeout_s);
        for (;;) {
            // Looked up again after every wake-up: a finished request can
            // leave the history while this thread sleeps.
            auto it = all_.find(id);
            if (it == all_.end()) return SGEMM_STATUS_NOT_FOUND;
            if (it->second->status != SGEMM_STATUS_NOT_READY) return it->second->status;
            if (done_cv_.wait_until(lock, deadline) == std::cv_status::timeout) {
                it = all_.find(id);
                if (it == all_.end()) return SGEMM_STATUS_NOT_FOUND;
                return it->second->status;
            }
        }
    }

    int drain(double timeout_s) {
        std::unique_lock<std::mutex> lock(mu_);
        const auto deadline = Clock::now() + std::chrono::duration<double>(timeout_s);
        while (completed_ + failed_ < submitted_) {
            if (done_cv_.wait_until(lock, deadline) == std::cv_status::timeout &&
                completed_ + failed_ < submitted_)
                return SGEMM_STATUS_NOT_READY;
        }
        return failed_ ? SGEMM_STATUS_HIP_ERROR : SGEMM_STATUS_SUCCESS;
    }

    int latency(uint64_t id, double* total_ms, double* queue_ms) {
        std::lock_guard<std::mutex> lock(mu_);
        auto it = all_.find(id);
        if (it == all_.end()) return SGEMM_STATUS_NOT_FOUND;
        const Req* q = it->second.get();
        if (q->status != SGEMM_STATUS_SUCCESS) return q->status;
        if (total_ms) *total_ms = (q->t_done - q->t_submit) * 1e3;
        if (queue_ms) *queue_ms = (q->t_admit - q->t_submit) * 1e3;
        return SGEMM_STATUS_SUCCESS;
    }

    // Running figures since the last reset.  Counts, means, the maximum and
    // the throughput cover every request measured since then; p50 and p99
    // the most recent kLatencyWindow of them, so the cost is bounded
    // whatever the server's history.
    void stats(sgemm_server_stats* out) {
        std::memset(out, 0, sizeof(*out));
        std::vector<double> lat;
        double qsum, first, last, flops, lsum, lmax;
        long long measured;
        {
            std::lock_guard<std::mutex> lock(mu_);
            out->submitted = submitted_;
            out->completed = completed_;
            out->rejected = rejected_;
            out->failed = failed_;
            out->batches = batches_;
            out->panel_launches = panel_launches_;
            out->mean_batch = batches_ ? (double)batched_requests_ / (double)batches_ : 0.0;
            out->peak_bytes = peak_bytes_;
            out->budget_bytes = cfg_.hbm_budget_bytes;
            lat = lat_ms_;
            qsum = queue_ms_sum_;
            first = first_submit_;
            last = last_done_;
            flops = flops_sum_;
            lsum = lat_sum_;
            lmax = lat_max_;
            measured = measured_;
        }
        out->measured = measured;
        if (lat.empty()) return;
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double p) {
            // nearest-rank percentile
            size_t r = (size_t)std::ceil(p / 100.0 * (double)lat.size());
            return lat[std::min(lat.size(), std::max<size_t>(r, 1)) - 1];
        };
        out->p50_ms = pct(50);
        out->p99_ms = pct(99);
        out->max_ms = lmax;
        out->mean_ms = lsum / (double)measured;
        out->mean_queue_ms = qsum / (double)measured;
        out->window_s = last - first;
        out->tflops = last > first ? flops / (last - first) / 1e12 : 0.0;
    }

    void reset_stats() {
        std::lock_guard<std::mutex> lock(mu_);
        stats_epoch_ = seconds_since(t0_);
        peak_bytes_ = used_bytes_;
        batches_ = 0;
        batched_requests_ = 0;
        panel_launches_ = 0;
        lat_ms_.clear();
        lat_next_ = 0;
        measured_ = 0;
        lat_sum_ = 0;
        lat_max_ = 0;
        queue_ms_sum_ = 0;
        flops_sum_ = 0;
        first_submit_ = 1e300;
        last_done_ = 0;
    }

  private:
    static size_t footprint(const sgemm_request& r) {
        return 4 * ((size_t)r.m * r.k + (size_t)r.k * r.n + (size_t)r.m * r.n);
    }

    int ensure_pool() {
        if (cfg_.executor != SGEMM_EXECUTOR_GPU) return SGEMM_STATUS_SUCCESS;  // offsets only
        std::lock_guard<std::mutex> lock(pool_mu_);
        if (pool_) return SGEMM_STATUS_SUCCESS;
        if (hipMalloc(&pool_, std::max<size_t>(cfg_.hbm_budget_bytes, 256)) != hipSuccess) {
            (void)hipGetLastError();
            pool_ = nullptr;
            return SGEMM_STATUS_ALLOC_FAILED;
        }
        return SGEMM_STATUS_SUCCESS;
    }

    // Scheduler thread.  Completion checks and launches run without mu_:
    // event queries, descriptor uploads, kernel launches and pageable copies
    // can take a while, and submit(), poll() and stats() must not wait
    // behind them.  What they touch is the scheduler's own (its slots) or
    // immutable while a request is in flight (its operands, panel layout and
    // pool offsets), and in-flight requests never leave all_.
    // How long the scheduler keeps polling after its last completion or
    // admission before it sleeps (see loop()).
    static constexpr double kIdleSpinS = 200e-6;

    void loop() {
        std::unique_lock<std::mutex> lock(mu_);
        std::vector<Slot*> planned;
        double last_work = seconds_since(t0_);
        while (!stop_) {
            lock.unlock();
            const bool any_done = poll_slots();
            lock.lock();
            if (any_done) retire();
            planned.clear();
            admit(&planned);
            if (any_done || !planned.empty()) last_work = seconds_since(t0_);
            if (!planned.empty()) {
                lock.unlock();
                for (Slot* s : planned) {
                    s->launch_status = s->chunk ? launch_chunk(*s) : launch(*s);
                    if (s->launch_status != SGEMM_STATUS_SUCCESS) {
                        // The slot retires once whatever was enqueued before
                        // the failure has drained; if even this record
                        // fails, the event still holds an earlier (completed)
                        // record.
                        s->mock_done_at = 0;
                        if (cfg_.executor == SGEMM_EXECUTOR_GPU)
                            (void)hipEventRecord(s->done, s->stream);
                    }
                }
                lock.lock();
                for (Slot* s : planned)
                    if (s->launch_status != SGEMM_STATUS_SUCCESS) launch_failed(*s);
                continue;  // more may be admissible now
            }
            // While a launch is in flight, and for kIdleSpinS after the last
            // completion or admission, the thread polls (yielding the core
            // between rounds) instead of sleeping: a timed wait returns only
            // after the kernel's timer slack, tens of microseconds, which
            // small requests would see in their latency (p50 at a quarter of
            // capacity 0.107 -> 0.076 ms).  Past that it sleeps until
            // submit() wakes it.
            bool any_busy = false;
            for (const Slot& s : slots_) any_busy |= s.busy;
            if (any_busy || seconds_since(t0_) - last_work < kIdleSpinS) {
                lock.unlock();
                std::this_thread::yield();
                lock.lock();
            } else {
                cv_.wait_for(lock, std::chrono::milliseconds(50));
            }
        }
    }

    // Marks busy slots whose launch has completed (no lock held).
    bool poll_slots() {
        bool any = false;
        const bool mock = cfg_.executor == SGEMM_EXECUTOR_MOCK;
        const double now = mock ? seconds_since(t0_) : 0.0;
        for (Slot& s : slots_) {
            if (!s.busy || s.finished) continue;
            if (mock) {
                if (now < s.mock_done_at) continue;
            } else {
                const hipError_t e = hipEventQuery(s.done);
                if (e == hipErrorNotReady) continue;
                if (e != hipSuccess) {
                    (void)hipGetLastError();
                    s.failed = true;
                }
            }
            s.finished = true;
            any = true;
        }
        return any;
    }

    // A launch returned an error (mu_ held): its requests fail once the slot
    // retires.  For a panel, the request's panels that were never launched
    // are dropped (counted as done), so the request finishes -- with the
    // error -- when its panels already in flight on other slots have retired,
    // not before: until then the caller must not free its operands.
    void launch_failed(Slot& s) {
        const int err = s.launch_status;
        if (Req* q = s.chunk) {
            if (q->error == SGEMM_STATUS_SUCCESS) q->error = err;
            q->panels_done += q->panels - q->next_panel;
            q->next_panel = q->panels;
            return;
        }
        for (Req* q : s.batch)
            if (q->error == SGEMM_STATUS_SUCCESS) q->error = err;
    }

    void finish(Req* q, double now) {
        q->t_done = now;
        q->status = q->error == SGEMM_STATUS_SUCCESS ? SGEMM_STATUS_SUCCESS : q->error;
        if (q->status == SGEMM_STATUS_SUCCESS) {
            ++completed_;
            if (q->t_submit >= stats_epoch_) {
                const double ms = (q->t_done - q->t_submit) * 1e3;
                if (lat_ms_.size() < kLatencyWindow) lat_ms_.push_back(ms);
                else lat_ms_[lat_next_] = ms;
                lat_next_ = (lat_next_ + 1) % kLatencyWindow;
                ++measured_;
                lat_sum_ += ms;
                lat_max_ = std::max(lat_max_, ms);
                queue_ms_sum_ += (q->t_admit - q->t_submit) * 1e3;
                flops_sum_ += 2.0 * q->r.m * q->r.n * (double)q->r.k;
                first_submit_ = std::min(first_submit_, q->t_submit);
                last_done_ = std::max(last_done_, q->t_done);
            }
        } else {
            ++failed_;
        }
        release(q);
        finished_ids_.push_back(q->id);
        while (finished_ids_.size() > kHistory) {
            all_.erase(finished_ids_.front());
            finished_ids_.pop_front();
        }
    }

    void retire() {
        bool any = false;
        const double now = seconds_since(t0_);
        for (Slot& s : slots_) {
            if (!s.busy || !s.finished) continue;
            if (Req* q = s.chunk) {
                ++q->panels_done;
                if (s.failed && q->error == SGEMM_STATUS_SUCCESS) q->error = SGEMM_STATUS_HIP_ERROR;
                if (q->panels_done == q->panels) finish(q, now);
                s.chunk = nullptr;
                s.chunk_panel = -1;
            } else {
                for (Req* q : s.batch) {
                    if (s.failed && q->error == SGEMM_STATUS_SUCCESS)
                        q->error = SGEMM_STATUS_HIP_ERROR;
                    finish(q, now);
                }
                s.batch.clear();
            }
            s.busy = s.finished = s.failed = false;
            any = true;
        }
        if (any) done_cv_.notify_all();
    }

    void release(Req* q) {
        if (q->pooled)
            for (size_t o : q->off) alloc_.release(o);
        q->pooled = false;
        used_bytes_ -= q->footprint;
    }

    // Reserve a request's footprint (and pool space for host operands),
    // keeping `headroom` bytes of the budget free.
    bool reserve(Req* q, size_t headroom = 0) {
        if (used_bytes_ + q->footprint + headroom > cfg_.hbm_budget_bytes) return false;
        if (q->r.operands == SGEMM_OPERANDS_HOST) {
            const size_t sz[3] = {4 * (size_t)q->r.m * q->r.k, 4 * (size_t)q->r.k * q->r.n,
                                  4 * (size_t)q->r.m * q->r.n};
            int got = 0;
            for (; got < 3; ++got)
                if (!alloc_.alloc(sz[got], &q->off[got])) break;
            if (got < 3) {
                for (int i = 0; i < got; ++i) alloc_.release(q->off[i]);
                return false;
            }
            q->pooled = true;
        }
        used_bytes_ += q->footprint;
        peak_bytes_ = std::max(peak_bytes_, used_bytes_);
        return true;
    }

    // Picks work for the free slots (mu_ held); the launches follow without
    // the lock.
    void admit(std::vector<Slot*>* planned) {
        for (Slot& s : slots_) {
            if (s.busy) continue;
            const bool large_first = prefer_large_;
            if (!(large_first ? plan_panel(s) || plan_small(s)
                              : plan_small(s) || plan_panel(s)))
                return;  // nothing admissible for any slot
            planned->push_back(&s);
        }
    }

    // Small lane: a grouped launch of up to max_batch requests from the head.
    bool plan_small(Slot& s) {
        // A large request waiting for memory keeps its footprint free: small
        // requests may overtake it, but never push it out of the budget.
        const size_t headroom =
            (!large_active_ && !large_q_.empty()) ? large_q_.front()->footprint : 0;
        std::vector<Req*> batch;
        while (!pending_.empty() && (int)batch.size() < cfg_.max_batch) {
            Req* q = pending_.front();
            if (!reserve(q, headroom)) break;  // FIFO within the lane
            q->t_admit = seconds_since(t0_);
            pending_.pop_front();
            batch.push_back(q);
        }
        if (batch.empty()) return false;
        s.batch = std::move(batch);
        s.busy = true;
        ++batches_;
        batched_requests_ += (long long)s.batch.size();
        prefer_large_ = true;
        return true;
    }

    // Large lane: the next panel of the active large request, admitting the
    // next one when the active request has launched all of its panels.
    bool plan_panel(Slot& s) {
        if (large_active_ && large_active_->next_panel == large_active_->panels)
            large_active_ = nullptr;
        if (!large_active_) {
            if (large_q_.empty()) return false;
            Req* q = large_q_.front();
            if (!reserve(q)) return false;
            q->t_admit = seconds_since(t0_);
            large_q_.pop_front();
            large_active_ = q;
        }
        Req* q = large_active_;
        s.chunk = q;
        s.chunk_panel = q->next_panel++;
        s.busy = true;
        ++panel_launches_;
        prefer_large_ = false;
        return true;
    }

    // Mock executor: the launch with (1-based) index cfg_.mock_fail_launch
    // fails, to exercise the failure paths without a GPU.
    bool mock_fails() {
        return cfg_.mock_fail_launch > 0 && ++mock_launches_ == cfg_.mock_fail_launch;
    }

    // One column panel as a single-problem launch (best config for its shape).
    int launch_chunk(Slot& s) {
        const sgemm_request& r = s.chunk->r;
        const int j0 = s.chunk_panel * s.chunk->panel_n;
        const int nc = std::min(s.chunk->panel_n, r.n - j0);
        if (cfg_.executor == SGEMM_EXECUTOR_MOCK) {
            if (mock_fails()) return SGEMM_STATUS_HIP_ERROR;
            s.mock_done_at = seconds_since(t0_) + cfg_.mock_launch_us * 1e-6 +
                             2.0 * r.m * nc * (double)r.k / (cfg_.mock_tflops * 1e12);
            return SGEMM_STATUS_SUCCESS;
        }
        const float* B = r.B + (size_t)j0 * r.ldb;
        float* C = r.C + (size_t)j0 * r.ldc;
        // The tail split and stream-K spread a partial round of tiles over
        // every CU, which assumes the launch has the GPU to itself; here
        // other slots' work shares it.  Such picks run the same tile whole:
        // at the benchmark's mixture that raised the capacity from 6450 to
        // 8400 requests/s.  A packed BF16x3 pick runs config 2 too (same
        // bits): its planes would be device memory outside the budget.
        int cfg = -1, from_table = 0;
        int whole = -1;
        if (sgemm_autotune_lookup(cfg_.form, 0, r.m, nc, r.k, &cfg, &from_table) ==
            SGEMM_STATUS_SUCCESS)
            whole = whole_tile_twin(cfg_.form, cfg);
        int st;
        if (whole >= 0 && r.lda <= kMaxLd && r.ldb <= kMaxLd && r.ldc <= kMaxLd)
            st = sgemm_run_config(cfg_.form, whole, r.m, nc, r.k, &r.alpha, r.A, r.lda, B, r.ldb,
                                  &r.beta, C, r.ldc, s.stream);
        else
            st = cfg_.form == SGEMM_FORM_BF16X3
                     ? sgemm_bf16x3(r.m, nc, r.k, &r.alpha, r.A, r.lda, B, r.ldb, &r.beta, C,
                                    r.ldc, s.stream)
                     : sgemm_ex(r.m, nc, r.k, &r.alpha, r.A, r.lda, B, r.ldb, &r.beta, C, r.ldc,
                                s.stream);
        if (st != SGEMM_STATUS_SUCCESS) return st;
        return hipEventRecord(s.done, s.stream) == hipSuccess ? SGEMM_STATUS_SUCCESS
                                                              : SGEMM_STATUS_HIP_ERROR;
    }

    int launch(Slot& s) {
        double flops = 0;
        for (Req* q : s.batch) flops += 2.0 * q->r.m * q->r.n * (double)q->r.k;
        if (cfg_.executor == SGEMM_EXECUTOR_MOCK) {
            if (mock_fails()) return SGEMM_STATUS_HIP_ERROR;
            s.mock_done_at = seconds_since(t0_) + cfg_.mock_launch_us * 1e-6 +
                             flops / (cfg_.mock_tflops * 1e12);
            return SGEMM_STATUS_SUCCESS;
        }
        std::vector<sgemm_problem> probs;
        probs.reserve(s.batch.size());
        for (Req* q : s.batch) {
            const sgemm_request& r = q->r;
            sgemm_problem p{r.m, r.n, r.k, r.alpha, r.beta, r.A, r.lda, r.B, r.ldb, r.C, r.ldc};
            if (r.operands == SGEMM_OPERANDS_HOST) {
                char* base = static_cast<char*>(pool_);
                float* dA = reinterpret_cast<float*>(base + q->off[0]);
                float* dB = reinterpret_cast<float*>(base + q->off[1]);
                float* dC = reinterpret_cast<float*>(base + q->off[2]);
                // pack host operands (column-major, ld) into dense device copies
                if (hipMemcpy2DAsync(dA, 4 * (size_t)r.m, r.A, 4 * (size_t)r.lda,
                                     4 * (size_t)r.m, r.k, hipMemcpyHostToDevice,
                                     s.stream) != hipSuccess ||
                    hipMemcpy2DAsync(dB, 4 * (size_t)r.k, r.B, 4 * (size_t)r.ldb,
                                     4 * (size_t)r.k, r.n, hipMemcpyHostToDevice,
                                     s.stream) != hipSuccess)
                    return SGEMM_STATUS_HIP_ERROR;
                if (r.beta != 0.f &&
                    hipMemcpy2DAsync(dC, 4 * (size_t)r.m, r.C, 4 * (size_t)r.ldc,
                                     4 * (size_t)r.m, r.n, hipMemcpyHostToDevice,
                                     s.stream) != hipSuccess)
                    return SGEMM_STATUS_HIP_ERROR;
                p.A = dA;
                p.lda = std::max(r.m, 1);
                p.B = dB;
                p.ldb = std::max(r.k, 1);
                p.C = dC;
                p.ldc = std::max(r.m, 1);
            }
            probs.push_back(p);
        }
        int st = s.plan->set((int)probs.size(), probs.data(), s.stream, nullptr);
        if (st == SGEMM_STATUS_SUCCESS) st = s.plan->run(s.stream);
        if (st != SGEMM_STATUS_SUCCESS) return st;
        for (size_t i = 0; i < s.batch.size(); ++i) {
            const Req* q = s.batch[i];
            const sgemm_request& r = q->r;
            if (r.operands != SGEMM_OPERANDS_HOST) continue;
            if (hipMemcpy2DAsync(r.C, 4 * (size_t)r.ldc, probs[i].C, 4 * (size_t)r.m,
                                 4 * (size_t)r.m, r.n, hipMemcpyDeviceToHost,
                                 s.stream) != hipSuccess)
                return SGEMM_STATUS_HIP_ERROR;
        }
        return hipEventRecord(s.done, s.stream) == hipSuccess ? SGEMM_STATUS_SUCCESS
                                                              : SGEMM_STATUS_HIP_ERROR;
    }

    sgemm_server_config cfg_;
    RangeAllocator alloc_;
    std::mutex pool_mu_;
    void* pool_ = nullptr;       // host-operand pool, allocated on first use
    std::vector<Slot> slots_;
    std::deque<Req*> pending_;   // small lane, arrival order
    std::deque<Req*> large_q_;   // large lane, arrival order
    Req* large_active_ = nullptr;
    bool prefer_large_ = false;
    double chunk_flops_ = 0;
    long long panel_launches_ = 0;
    int mock_launches_ = 0;      // scheduler thread only
    std::unordered_map<uint64_t, std::unique_ptr<Req>> all_;
    std::deque<uint64_t> finished_ids_;  // oldest first; evicted past kHistory
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::thread worker_;
    bool stop_ = false;
    Clock::time_point t0_;
    uint64_t next_id_ = 0;
    long long submitted_ = 0, completed_ = 0, rejected_ = 0, failed_ = 0;
    long long batches_ = 0, batched_requests_ = 0;
    size_t used_bytes_ = 0, peak_bytes_ = 0;
    double stats_epoch_ = 0;
    // Running statistics of the requests submitted since stats_epoch_:
    // the latency ring (the last kLatencyWindow, lat_next_ the slot written
    // next once it is full) and sums over all of them.
    std::vector<double> lat_ms_;
    size_t lat_next_ = 0;
    long long measured_ = 0;
    double lat_sum_ = 0, lat_max_ = 0;
    double queue_ms_sum_ = 0, flops_sum_ = 0, first_submit_ = 1e300, last_done_ = 0;
};

}  // namespace sgemm_amd

using sgemm_amd::Server;

extern "C" {

int sgemm_server_create(sgemm_server* srv, const sgemm_server_config* cfg) {
    if (srv == nullptr || cfg == nullptr) return SGEMM_STATUS_INVALID_VALUE;
    *srv = nullptr;
    if (cfg->hbm_budget_bytes == 0 || cfg->max_batch < 1 || cfg->num_slots < 1 ||
        cfg->num_slots > 16)
        return SGEMM_STATUS_INVALID_VALUE;
    if (cfg->form != SGEMM_FORM_F32 && cfg->form != SGEMM_FORM_BF16X3)
        return SGEMM_STATUS_NOT_SUPPORTED;
    if (cfg->executor != SGEMM_EXECUTOR_GPU && cfg->executor != SGEMM_EXECUTOR_MOCK)
        return SGEMM_STATUS_INVALID_VALUE;
    if (cfg->executor == SGEMM_EXECUTOR_MOCK && !(cfg->mock_tflops > 0))
        return SGEMM_STATUS_INVALID_VALUE;
    Server* s = new Server(*cfg);
    int st = s->start();
    if (st != SGEMM_STATUS_SUCCESS) {
        delete s;
        return st;
    }
    *srv = reinterpret_cast<sgemm_server>(s);
    return SGEMM_STATUS_SUCCESS;
}

int sgemm_server_submit(sgemm_server srv, const sgemm_request* req, uint64_t* id) {
    if (srv == nullptr || req == nullptr) return SGEMM_STATUS_INVALID_VALUE;
    return reinterpret_cast<Server*>(srv)->submit(*req, id);
}

int sgemm_server_poll(sgemm_server srv, uint64_t id) {
    if (srv == nullptr) return SGEMM_STATUS_INVALID_VALUE;
    return reinterpret_cast<Server*>(srv)->poll(id);
}

int sgemm_server_wait(sgemm_server srv, uint64_t id, double timeout_s) {
    if (srv == nullptr) return SGEMM_STATUS_INVALID_VALUE;
    return reinterpret_cast<Server*>(srv)->wait(id, timeout_s);
}

int sgemm_server_drain(sgemm_server srv, double timeout_s) {
    if (srv == nullptr) return SGEMM_STATUS_INVALID_VALUE;
    return reinterpret_cast<Server*>(srv)->drain(timeout_s);
}

int sgemm_server_latency(sgemm_server srv, uint64_t id, double* total_ms, double* queue_ms) {
    if (srv == nullptr) return SGEMM_STATUS_INVALID_VALUE;
    return reinterpret_cast<Server*>(srv)->latency(id, total_ms, queue_ms);
}

int sgemm_server_get_stats(sgemm_server srv, sgemm_server_stats* out) {
    if (srv == nullptr || out == nullptr) return SGEMM_STATUS_INVALID_VALUE;
    reinterpret_cast<Server*>(srv)->stats(out);
    return SGEMM_STATUS_SUCCESS;
}

int sgemm_server_reset_stats(sgemm_server srv) {
    if (srv == nullptr) return SGEMM_STATUS_INVALID_VALUE;
    reinterpret_cast<Server*>(srv)->reset_stats();
    return SGEMM_STATUS_SUCCESS;
}

int sgemm_server_destroy(sgemm_server srv) {
    if (srv == nullptr) return SGEMM_STATUS_INVALID_VALUE;
    delete reinterpret_cast<Server*>(srv);
    return SGEMM_STATUS_SUCCESS;
}

}  // extern "C"
