"""GPU parity at the benchmark's own scale.

* Grouped (512 problems): exactly the benchmark's group -- M, N, K
  log-uniform integers in [32, 2048] from seed 0, alpha = 1, beta = 0 --
  in both forms, with every size bucket tuned by the autotuner, run as a
  captured hipGraph that is replayed.  Every problem must equal a single
  launch of the same config bit for bit, and meet the FP64 oracle on sampled
  entries.
* Serving: the benchmark's server (64 GB budget, max batch 32, 3 slots,
  default 100 GFLOP panels) on a mix with requests from all three size
  buckets, three of them from [4096, 8192]: each result is the same bits
  alone and in the mix, and meets the FP64 oracle on sampled entries.
* Panels whose best config is the tail split or stream-K run that config's
  tile whole: the result equals a single launch of the whole-tile config.
"""
import time

import numpy as np
import pytest

import oracle as orc
import sgemm_amd as sa
from gpu_util import Case, download, log, upload

pytestmark = pytest.mark.gpu

GB = 1 << 30


def _tol(form):
    return orc.fp32_tol_mult if form == sa.FORM_F32 else orc.bf16x3_tol_mult


def _sampled_ok(c, out, form, n_samples, seed):
    """FP64 oracle on sampled entries (the corners included) of an
    ldc-strided result."""
    rng = np.random.default_rng(seed)
    ii = np.concatenate([rng.integers(0, c.m, n_samples), [0, c.m - 1]]).astype(np.int32)
    jj = np.concatenate([rng.integers(0, c.n, n_samples), [c.n - 1, 0]]).astype(np.int32)
    ref, scale = orc.dgemm_entries(c.k, c.alpha, c.A, c.lda, c.B, c.ldb, c.beta, c.C0, c.ldc,
                                   ii, jj)
    got = out[ii + jj.astype(np.int64) * c.ldc].astype(np.float64)
    err = np.abs(got - ref)
    tol = _tol(form)(c.k) * scale
    assert np.all(np.isfinite(got)) and np.all(err <= tol), \
        ((c.m, c.n, c.k), float(np.max(err / tol)))


def _bench_group_shapes(count=512, seed=0):
    # the benchmark's draw (bench.py grouped_leg)
    rng = np.random.default_rng(seed)
    v = np.floor(np.exp(rng.uniform(np.log(32), np.log(2049), size=(count, 3)))).astype(int)
    return [tuple(int(x) for x in row) for row in v]


@pytest.mark.parametrize("form", [sa.FORM_F32, sa.FORM_BF16X3])
def test_grouped_512_tuned_captured(form):
    import torch
    shapes = _bench_group_shapes()
    cases, dev, probs = [], [], []
    for i, (m, n, k) in enumerate(shapes):
        c = Case(m, n, k, seed=9000 + 7 * i, alpha=1.0, beta=0.0)
        dA, dB, dC = c.device()
        cases.append(c)
        dev.append((dA, dB, dC))
        probs.append(sa.problem(m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m))
    sa.autotune_clear()
    stream = torch.cuda.Stream()
    try:
        t0 = time.perf_counter()
        tuned = {}
        for (m, n, k) in shapes:
            key = tuple(int(v - 1).bit_length() for v in (m, n, k))
            if key not in tuned:
                tuned[key] = sa.autotune_bucket(form, m, n, k, iters=2, grouped=True,
                                                stream=stream)[0]
        log(f"grouped 512 form {form}: {len(tuned)} buckets tuned in "
            f"{time.perf_counter() - t0:.1f} s")
        configs = [sa.autotune_lookup(form, m, n, k, grouped=True) for (m, n, k) in shapes]
        assert all(from_table for _, from_table in configs)
        plan = sa.GroupedPlan(len(probs), form)
        with torch.cuda.stream(stream):
            plan.set(probs, stream)
            plan.capture(stream)
            for rep in range(2):
                for _, _, dC in dev:
                    dC.fill_(np.nan)  # every replay must write every entry
                stream.synchronize()
                log(f"grouped 512 form {form}: replay {rep}")
                plan.replay(stream)
                stream.synchronize()
        info = plan.info()
        assert info["count"] == 512 and info["total_tiles"] > 512
        for i, (c, (dA, dB, dC), (cfg, _)) in enumerate(zip(cases, dev, configs)):
            got = download(dC)
            single = upload(np.zeros(c.m * c.n, np.float32))
            sa.run_config(form, cfg, c.m, c.n, c.k, 1.0, dA, c.m, dB, c.k, 0.0, single, c.m)
            assert np.array_equal(got, download(single)), (i, (c.m, c.n, c.k), cfg)
            _sampled_ok(c, got, form, 16, seed=i)
        plan.destroy()
    finally:
        sa.autotune_clear()


def _mixture(seed, counts=(20, 7, 3)):
    """Requests from the benchmark's three size buckets ([128, 512],
    [1024, 2048], [4096, 8192]), shuffled."""
    rng = np.random.default_rng(seed)
    sizes = []
    for cnt, (lo, hi) in zip(counts, ((128, 512), (1024, 2048), (4096, 8192))):
        sizes += [tuple(int(v) for v in rng.integers(lo, hi + 1, 3)) for _ in range(cnt)]
    order = rng.permutation(len(sizes))
    return [sizes[i] for i in order]


@pytest.mark.parametrize("form", [sa.FORM_BF16X3, sa.FORM_F32])
def test_serving_bench_config_alone_equals_mix(form):
    import torch
    sizes = _mixture(seed=77 + form)
    cases = [Case(m, n, k, seed=5000 + 11 * i, alpha=1.0, beta=0.5)
             for i, (m, n, k) in enumerate(sizes)]
    dev = [(upload(c.A), upload(c.B)) for c in cases]
    srv = sa.Server(64 * GB, max_batch=32, num_slots=3, form=form)
    alone = []
    for c, (dA, dB) in zip(cases, dev):
        out = upload(c.C0)
        log(f"serve bench-config alone form {form} {(c.m, c.n, c.k)}")
        rid = srv.submit(c.m, c.n, c.k, c.alpha, dA, c.lda, dB, c.ldb, c.beta, out, c.ldc)
        assert srv.wait(rid, 120) == 0
        alone.append(download(out))
        del out
    outs = [upload(c.C0) for c in cases]
    log(f"serve bench-config mix form {form}")
    ids = [srv.submit(c.m, c.n, c.k, c.alpha, dA, c.lda, dB, c.ldb, c.beta, o, c.ldc)
           for c, (dA, dB), o in zip(cases, dev, outs)]
    srv.drain(300)
    st = srv.stats()
    assert st["failed"] == 0 and st["completed"] == 2 * len(cases)
    assert st["panel_launches"] > 0 and st["peak_bytes"] <= 64 * GB
    for i, (c, a, o, rid) in enumerate(zip(cases, alone, outs, ids)):
        assert srv.poll(rid) == 0
        mixed = download(o)
        assert np.array_equal(a, mixed), (c.m, c.n, c.k)
        _sampled_ok(c, mixed, form, 256, seed=i)
    srv.destroy()
    torch.cuda.synchronize()


@pytest.mark.parametrize("form,split_cfg", [(sa.FORM_BF16X3, 11), (sa.FORM_BF16X3, 9),
                                            (sa.FORM_F32, 11), (sa.FORM_F32, 9),
                                            (sa.FORM_F32, 16)])
def test_panels_run_split_picks_as_whole_tiles(form, split_cfg):
    """With the panel shapes' buckets set to the stream-K (11, FP32 16) or
    tail-split (9) config, the server runs each panel on the same tile whole
    (BF16x3 config 2, FP32 config 5 or 15, which gives config 5's bits): the
    request's result equals one launch of config 2 or 5 over the full
    problem, alone and in a mix."""
    import torch
    whole = 2 if form == sa.FORM_BF16X3 else 5
    m, n, k = 4096, 5000, 4096            # 168 GFLOP: panels of 2816 and 2184 columns
    panel_n = (int(100e9 / (2.0 * m * k)) // 256) * 256
    widths = [panel_n, n - panel_n]
    sa.autotune_clear()
    try:
        for w in widths:
            sa.autotune_set(form, m, w, k, split_cfg)
            assert sa.autotune_lookup(form, m, w, k) == (split_cfg, True)
        cases = [Case(m, n, k, seed=600 + s, alpha=1.0, beta=0.5) for s in range(2)]
        dev = [c.device() for c in cases]
        srv = sa.Server(64 * GB, max_batch=32, num_slots=3, form=form)
        log(f"serve split pick {split_cfg} form {form}: alone")
        rid = srv.submit(m, n, k, 1.0, dev[0][0], m, dev[0][1], k, 0.5, dev[0][2], m)
        assert srv.wait(rid, 60) == 0
        alone = download(dev[0][2])
        want = upload(cases[0].C0)
        sa.run_config(form, whole, m, n, k, 1.0, dev[0][0], m, dev[0][1], k, 0.5, want, m)
        assert np.array_equal(alone, download(want))
        # both requests at once, plus small traffic between their panels
        outs = [upload(c.C0) for c in cases]
        log(f"serve split pick {split_cfg} form {form}: mix")
        ids = [srv.submit(m, n, k, 1.0, d[0], m, d[1], k, 0.5, o, m)
               for d, o in zip(dev, outs)]
        small = Case(300, 200, 500, seed=9, alpha=1.0, beta=0.5)
        sA, sB, sC = small.device()
        ids.append(srv.submit(300, 200, 500, 1.0, sA, 300, sB, 500, 0.5, sC, 300))
        srv.drain(120)
        assert srv.stats()["failed"] == 0
        assert np.array_equal(download(outs[0]), alone)
        for i, (c, o) in enumerate(zip(cases, outs)):
            _sampled_ok(c, download(o), form, 512, seed=i)
        _sampled_ok(small, download(sC), form, 512, seed=3)
        srv.destroy()
        torch.cuda.synchronize()
    finally:
        sa.autotune_clear()


def test_device_operand_server_allocates_no_pool():
    """The host-operand pool (the whole budget) is allocated by the first
    host-operand request, not when the server is created."""
    import torch
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    srv = sa.Server(48 * GB, max_batch=8, num_slots=2, form=sa.FORM_BF16X3)
    c = Case(256, 200, 300, seed=4)
    dA, dB, dC = c.device()
    rid = srv.submit(c.m, c.n, c.k, c.alpha, dA, c.lda, dB, c.ldb, c.beta, dC, c.ldc)
    assert srv.wait(rid, 30) == 0
    free1, _ = torch.cuda.mem_get_info()
    assert free0 - free1 < 4 * GB, (free0 - free1) / GB
    host_C = c.C0.copy()
    rid = srv.submit(c.m, c.n, c.k, c.alpha, c.A, c.lda, c.B, c.ldb, c.beta, host_C, c.ldc,
                     operands=sa.OPERANDS_HOST)
    assert srv.wait(rid, 30) == 0
    free2, _ = torch.cuda.mem_get_info()
    assert free1 - free2 > 40 * GB, (free1 - free2) / GB
    assert np.array_equal(host_C, download(dC))
    srv.destroy()
