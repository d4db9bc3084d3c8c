"""Serving-loop scheduling on the mock executor (no GPU): admission against
the HBM budget, strict FIFO, continuous batching into launches, slot reuse,
rejection of requests that can never fit, and the latency statistics.  The
mock completes a launch after flops / mock_tflops seconds of wall time."""
import time

import numpy as np
import pytest

import sgemm_amd as sa

MB = 1 << 20


@pytest.fixture(autouse=True)
def built():
    import os
    if not os.path.exists(sa.LIB_PATH):
        sa.build()


def _submit(srv, m, n, k, **kw):
    # device-operand requests with dummy pointers: the mock never dereferences
    return srv.submit(m, n, k, 1.0, 4096, m, 4096, k, 0.0, 4096, m, **kw)


def test_config_validation():
    for bad in [dict(budget_bytes=0), dict(budget_bytes=MB, max_batch=0),
                dict(budget_bytes=MB, num_slots=0), dict(budget_bytes=MB, num_slots=17),
                dict(budget_bytes=MB, form=5), dict(budget_bytes=MB, mock_tflops=0.0)]:
        kw = dict(budget_bytes=MB, executor=sa.EXECUTOR_MOCK, mock_tflops=1.0)
        kw.update(bad)
        with pytest.raises(sa.SgemmError):
            sa.Server(**kw)


def test_all_requests_complete_and_stats():
    srv = sa.Server(64 * MB, max_batch=8, num_slots=2, executor=sa.EXECUTOR_MOCK,
                    mock_tflops=0.05, mock_launch_us=200)
    rng = np.random.default_rng(0)
    ids = []
    for _ in range(40):
        m, n, k = (int(v) for v in rng.integers(64, 256, 3))
        ids.append(_submit(srv, m, n, k))
    srv.drain(30)
    st = srv.stats()
    assert st["submitted"] == 40 and st["completed"] == 40 and st["failed"] == 0
    assert st["measured"] == 40 and 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert st["batches"] >= 40 / 8 and st["mean_batch"] > 1.0  # batched, not one by one
    assert st["tflops"] > 0 and st["peak_bytes"] <= 64 * MB
    for rid in ids:
        assert srv.poll(rid) == 0
        total, queue = srv.latency(rid)
        assert total >= queue >= 0
    assert srv.poll(10 ** 9) == 7  # unknown id
    srv.destroy()


def test_budget_admission_and_rejection():
    # 1 MiB budget; one 256x256x256 request needs 3 * 256 KiB = 768 KiB
    srv = sa.Server(1 * MB, max_batch=8, num_slots=4, executor=sa.EXECUTOR_MOCK,
                    mock_tflops=0.001, mock_launch_us=0)
    too_big = _submit(srv, 512, 512, 512, check=False)
    assert too_big == -5  # BUDGET_EXCEEDED
    a = _submit(srv, 256, 256, 256)
    b = _submit(srv, 256, 256, 256)
    time.sleep(0.005)
    # only one fits: the second waits although three slots are idle
    assert srv.poll(a) == 6 and srv.poll(b) == 6
    st = srv.stats()
    assert st["rejected"] == 1 and st["peak_bytes"] <= 1 * MB
    srv.drain(30)
    ta, qa = srv.latency(a)
    tb, qb = srv.latency(b)
    # b was admitted only after a finished and released its footprint
    assert qb >= ta - 1.0
    srv.destroy()


def test_strict_fifo_no_overtaking():
    srv = sa.Server(1 * MB, max_batch=8, num_slots=4, executor=sa.EXECUTOR_MOCK,
                    mock_tflops=0.002, mock_launch_us=0)
    big1 = _submit(srv, 256, 256, 256)   # 768 KiB
    big2 = _submit(srv, 256, 256, 256)   # must wait for big1
    small = _submit(srv, 16, 16, 16)     # fits beside big1, but may not overtake big2
    srv.drain(30)
    _, q_big2 = srv.latency(big2)
    _, q_small = srv.latency(small)
    assert q_small >= q_big2 - 0.5
    srv.destroy()


def test_continuous_batching_admits_while_busy():
    """Requests arriving while a launch is in flight go out in the next free
    slot without waiting for the first launch to finish."""
    srv = sa.Server(64 * MB, max_batch=4, num_slots=2, executor=sa.EXECUTOR_MOCK,
                    mock_tflops=0.0005, mock_launch_us=0)
    first = _submit(srv, 256, 256, 256)      # ~67 ms of mock compute
    time.sleep(0.01)
    second = _submit(srv, 32, 32, 32)        # tiny
    srv.wait(second, 10)
    assert srv.poll(first) == 6              # first still running
    srv.drain(30)
    t2, q2 = srv.latency(second)
    t1, _ = srv.latency(first)
    assert q2 < 5.0 and t2 < t1
    srv.destroy()


def test_host_operand_pool_fragmentation_is_safe():
    """Host-operand requests carve A, B, C out of the pool; mixed sizes must
    all complete and never exceed the pool."""
    srv = sa.Server(2 * MB, max_batch=3, num_slots=3, executor=sa.EXECUTOR_MOCK,
                    mock_tflops=0.01, mock_launch_us=50)
    rng = np.random.default_rng(4)
    bufs = []
    for _ in range(30):
        m, n, k = (int(v) for v in rng.integers(8, 200, 3))
        A = np.zeros(m * k, np.float32)
        B = np.zeros(k * n, np.float32)
        C = np.zeros(m * n, np.float32)
        bufs.append((A, B, C))
        srv.submit(m, n, k, 1.0, A, m, B, k, 0.0, C, m, operands=sa.OPERANDS_HOST)
    srv.drain(30)
    st = srv.stats()
    assert st["completed"] == 30 and st["peak_bytes"] <= 2 * MB
    srv.destroy()


def test_invalid_request_rejected():
    srv = sa.Server(MB, executor=sa.EXECUTOR_MOCK, mock_tflops=1.0)
    assert srv.submit(4, 4, 4, 1.0, 4096, 3, 4096, 4, 0.0, 4096, 4, check=False) == -1
    assert srv.submit(4, 4, 4, 1.0, 4096, 4, 4096, 4, 0.0, 4096, 4, operands=7,
                      check=False) == -1
    srv.destroy()


def test_large_request_is_cut_into_panels_and_small_ones_overtake():
    """A request above the chunk threshold runs as column panels (one launch
    each); small requests submitted behind it are interleaved between the
    panels, so they finish long before it does."""
    srv = sa.Server(256 * MB, max_batch=8, num_slots=1, executor=sa.EXECUTOR_MOCK,
                    mock_tflops=0.01, mock_launch_us=0, chunk_gflop=0.01)
    large = _submit(srv, 512, 1024, 512)        # 537 MFLOP -> 4 panels of 256 columns
    time.sleep(0.002)
    small = [_submit(srv, 64, 64, 64) for _ in range(6)]
    srv.drain(30)
    st = srv.stats()
    assert st["completed"] == 7 and st["panel_launches"] == 4
    t_large, _ = srv.latency(large)
    for rid in small:
        t, _ = srv.latency(rid)
        # at most about one panel (13.4 ms of mock compute) ahead of each
        assert t < 0.6 * t_large, (t, t_large)
    srv.destroy()


def test_large_request_does_not_starve_under_small_traffic():
    """Panels alternate with small batches, so a large request keeps advancing
    while small requests keep arriving."""
    srv = sa.Server(256 * MB, max_batch=4, num_slots=1, executor=sa.EXECUTOR_MOCK,
                    mock_tflops=0.01, mock_launch_us=0, chunk_gflop=0.01)
    large = _submit(srv, 512, 1024, 512)        # 4 panels, ~13.4 ms each
    for _ in range(60):                          # ~120 ms of steady small traffic
        _submit(srv, 128, 128, 128)              # 0.4 ms each
        time.sleep(0.002)
    srv.wait(large, 30)
    t_large, _ = srv.latency(large)
    # 4 panels plus at most one small batch (<= 4 x 0.4 ms) before each
    assert t_large < 4 * (13.4 + 1.6) + 40.0, t_large
    srv.drain(30)
    assert srv.stats()["completed"] == 61
    srv.destroy()


def test_small_requests_leave_room_for_a_waiting_large_one():
    """A large request waiting for memory keeps its footprint reserved against
    small overtakers: the small ones never push it out of the budget."""
    # footprints: 512x1024x512 -> 4*(256K + 512K + 512K) = 5 MiB; 256^3 -> 768 KiB
    budget = 5 * MB + MB // 2   # `first` + large does not fit; large alone does
    srv = sa.Server(budget, max_batch=8, num_slots=3, executor=sa.EXECUTOR_MOCK,
                    mock_tflops=0.002, mock_launch_us=0, chunk_gflop=0.01)
    first = _submit(srv, 256, 256, 256)          # holds 768 KiB for ~17 ms
    time.sleep(0.002)
    large = _submit(srv, 512, 1024, 512)         # 5 MiB: waits for `first`
    small = [_submit(srv, 256, 256, 256) for _ in range(3)]
    srv.drain(60)
    st = srv.stats()
    assert st["completed"] == 5 and st["peak_bytes"] <= budget
    _, q_large = srv.latency(large)
    t_first, _ = srv.latency(first)
    # the large request was admitted as soon as `first` released its memory,
    # not after the later small requests, which could not overtake it
    assert q_large < t_first + 5.0
    for rid in small:
        _, q = srv.latency(rid)
        assert q >= q_large - 0.5
    srv.destroy()


def test_panel_failure_is_reported_after_inflight_panels():
    """One panel of a large request fails at launch while another panel of
    it is still running on the other slot: the request's error is published
    only once that panel has retired (the caller may free the operands on
    the error), and its remaining panels are never launched."""
    srv = sa.Server(256 * MB, max_batch=8, num_slots=2, executor=sa.EXECUTOR_MOCK,
                    mock_tflops=0.01, mock_launch_us=0, chunk_gflop=0.01, mock_fail_launch=2)
    t0 = time.perf_counter()
    large = _submit(srv, 512, 1024, 512)   # 4 panels of ~13.4 ms; launch 2 fails
    time.sleep(0.003)
    assert srv.poll(large) == 6            # panel 0 still in flight: not reported yet
    assert srv.wait(large, 10) == 3        # SGEMM_STATUS_HIP_ERROR
    elapsed_ms = (time.perf_counter() - t0) * 1e3
    assert elapsed_ms >= 10.0, elapsed_ms  # waited for panel 0 (13.4 ms of mock compute)
    st = srv.stats()
    assert st["failed"] == 1 and st["completed"] == 0
    assert st["panel_launches"] == 2       # panels 2 and 3 were dropped
    # the server keeps serving afterwards
    ok = _submit(srv, 64, 64, 64)
    assert srv.wait(ok, 10) == 0
    with pytest.raises(sa.SgemmError):     # drain reports that a request failed
        srv.drain(10)
    srv.destroy()


def test_batch_failure_fails_only_its_requests():
    srv = sa.Server(64 * MB, max_batch=2, num_slots=1, executor=sa.EXECUTOR_MOCK,
                    mock_tflops=0.001, mock_launch_us=0, mock_fail_launch=1)
    ids = [_submit(srv, 64, 64, 64) for _ in range(6)]
    for rid in ids:
        srv.wait(rid, 10)
    codes = [srv.poll(rid) for rid in ids]
    assert codes.count(3) >= 1 and codes.count(3) <= 2, codes   # the first batch only
    assert codes.count(0) == len(ids) - codes.count(3), codes
    assert codes[-1] == 0
    st = srv.stats()
    assert st["failed"] == codes.count(3) and st["measured"] == codes.count(0)
    srv.destroy()


def test_history_is_bounded_and_stats_stay_complete():
    """Finished requests leave the history past 65536 (their ids return
    NOT_FOUND); the running statistics still cover every request (counts,
    mean and max over all of them, the percentiles over the most recent
    65536 latencies)."""
    srv = sa.Server(1 << 30, max_batch=256, num_slots=4, executor=sa.EXECUTOR_MOCK,
                    mock_tflops=1e6, mock_launch_us=0)
    count = 65536 + 1000
    ids = [_submit(srv, 8, 8, 8) for _ in range(count)]
    srv.drain(120)
    codes = [srv.poll(rid) for rid in ids]
    assert codes.count(7) == 1000 and codes.count(0) == 65536    # oldest 1000 evicted
    assert codes[0] == 7 and codes[-1] == 0
    st = srv.stats()
    assert st["completed"] == count and st["measured"] == count
    kept = [srv.latency(rid)[0] for rid in ids[-65536:]]
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert st["max_ms"] >= max(kept) and 0 < st["mean_ms"] <= st["max_ms"]
    srv.reset_stats()
    assert srv.stats()["measured"] == 0
    srv.destroy()
