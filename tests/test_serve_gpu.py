"""Serving loop on the GPU: every request's result is bit-identical alone and
in a mix of concurrent requests (same config per shape, same tile code), and
within the stated tolerance of the FP64 oracle; host-operand requests give the
same bits as device-operand ones."""
import time

import numpy as np
import pytest

import oracle as orc
import sgemm_amd as sa
from gpu_util import Case, check, download, log, upload

pytestmark = pytest.mark.gpu

GB = 1 << 30


def _cases(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        r = rng.random()
        lo, hi = (128, 512) if r < 0.7 else (1024, 2048) if r < 0.97 else (2048, 3072)
        m, n, k = (int(v) for v in rng.integers(lo, hi + 1, 3))
        out.append(Case(m, n, k, seed=seed * 1000 + 3 * i, alpha=1.0, beta=0.5))
    return out


@pytest.mark.parametrize("form,chunk_gflop", [(sa.FORM_BF16X3, 8.0), (sa.FORM_F32, 8.0),
                                               (sa.FORM_BF16X3, -1.0)])
def test_alone_equals_mix_and_oracle(form, chunk_gflop):
    """chunk_gflop = 8 cuts the requests above 8 GFLOP (most of the
    1024..3072 ones) into column panels run as single launches between the
    grouped batches; -1 sends everything through grouped batches."""
    import torch
    cases = _cases(24, seed=7 + form)
    dev = [c.device() for c in cases]
    srv = sa.Server(8 * GB, max_batch=8, num_slots=3, form=form, chunk_gflop=chunk_gflop)

    # alone: one request in the system at a time
    alone = []
    for c, (dA, dB, dC) in zip(cases, dev):
        out = upload(c.C0)
        log(f"serve alone form {form} chunk {chunk_gflop} {(c.m, c.n, c.k)}")
        rid = srv.submit(c.m, c.n, c.k, c.alpha, dA, c.lda, dB, c.ldb, c.beta, out, c.ldc)
        assert srv.wait(rid, 60) == 0
        alone.append(download(out))

    # mix: all at once, plus a second wave while the first is in flight
    outs = [upload(c.C0) for c in cases]
    ids = []
    log(f"serve mix form {form} chunk {chunk_gflop}")
    for i, (c, (dA, dB, _)) in enumerate(zip(cases, dev)):
        ids.append(srv.submit(c.m, c.n, c.k, c.alpha, dA, c.lda, dB, c.ldb, c.beta, outs[i],
                              c.ldc))
        if i == 12:
            time.sleep(0.002)
    srv.drain(120)
    st = srv.stats()
    assert st["failed"] == 0 and st["completed"] == 2 * len(cases)
    assert st["mean_batch"] > 1.0
    assert (st["panel_launches"] > 0) == (chunk_gflop > 0)
    tol = orc.fp32_tol_mult if form == sa.FORM_F32 else orc.bf16x3_tol_mult
    for c, a, o, rid in zip(cases, alone, outs, ids):
        assert srv.poll(rid) == 0
        mixed = download(o)
        assert np.array_equal(a, mixed), (c.m, c.n, c.k)
        s = check(c, mixed, tol(c.k))
        assert s["n_fail"] == 0, ((c.m, c.n, c.k), s)
    srv.destroy()
    torch.cuda.synchronize()


def test_host_operands_match_device_operands():
    cases = _cases(10, seed=21)
    srv = sa.Server(2 * GB, max_batch=4, num_slots=2, form=sa.FORM_BF16X3)
    for c in cases:
        dA, dB, dC = c.device()
        rid = srv.submit(c.m, c.n, c.k, c.alpha, dA, c.lda, dB, c.ldb, c.beta, dC, c.ldc)
        assert srv.wait(rid, 60) == 0
        host_C = c.C0.copy()
        log(f"serve host operands {(c.m, c.n, c.k)}")
        rid = srv.submit(c.m, c.n, c.k, c.alpha, c.A, c.lda, c.B, c.ldb, c.beta, host_C, c.ldc,
                         operands=sa.OPERANDS_HOST)
        assert srv.wait(rid, 60) == 0
        assert np.array_equal(host_C, download(dC)), (c.m, c.n, c.k)
    srv.destroy()


def test_budget_blocks_then_admits():
    """A budget that holds one large request at a time: both complete, and the
    second one is admitted only after the first has finished.  Both
    latencies count from each request's own submission, and the host gap
    between the two submissions bounds the difference:
    admit2 >= done1  <=>  q2 >= t1 - (submit2 - submit1) >= t1 - gap."""
    c1 = Case(2048, 2048, 2048, seed=31, alpha=1.0, beta=0.0)
    c2 = Case(2048, 2048, 2048, seed=32, alpha=1.0, beta=0.0)
    budget = 4 * 3 * 2048 * 2048 + (1 << 20)   # one request's footprint + 1 MiB
    srv = sa.Server(budget, max_batch=4, num_slots=2, form=sa.FORM_BF16X3)
    d1, d2 = c1.device(), c2.device()
    t0 = time.perf_counter()
    r1 = srv.submit(2048, 2048, 2048, 1.0, d1[0], 2048, d1[1], 2048, 0.0, d1[2], 2048)
    r2 = srv.submit(2048, 2048, 2048, 1.0, d2[0], 2048, d2[1], 2048, 0.0, d2[2], 2048)
    gap_ms = (time.perf_counter() - t0) * 1e3
    srv.drain(60)
    t1, _ = srv.latency(r1)
    _, q2 = srv.latency(r2)
    assert q2 >= t1 - gap_ms - 1e-3, (q2, t1, gap_ms)
    assert srv.stats()["peak_bytes"] <= budget
    for c, d in ((c1, d1), (c2, d2)):
        s = check(c, download(d[2]), orc.bf16x3_tol_mult(c.k))
        assert s["n_fail"] == 0, s
    srv.destroy()
