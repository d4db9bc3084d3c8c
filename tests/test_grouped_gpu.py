"""GPU tests of the grouped persistent launch, its hipGraph capture and the
autotune table.

The grouped result of every problem must be BIT-IDENTICAL to a single launch
of the same problem with the same config (same tile code, same summation
order), and within the stated tolerance of the FP64 oracle.
"""
import os
import tempfile

import numpy as np
import pytest

import oracle as orc
import sgemm_amd as sa
from gpu_util import Case, check, compact, download, log, upload

pytestmark = pytest.mark.gpu


def _log_uniform_shapes(count, lo=32, hi=2048, seed=0):
    rng = np.random.default_rng(seed)
    v = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi + 1), size=(count, 3)))).astype(int)
    return [tuple(int(x) for x in row) for row in v]


def _make_group(shapes, seed=0, kind="uniform", alpha=1.0, beta=0.0, ld_pad=0):
    cases, dev, probs = [], [], []
    for i, (m, n, k) in enumerate(shapes):
        c = Case(m, n, k, lda=max(1, m + ld_pad), ldb=max(1, k + ld_pad),
                 ldc=max(1, m + ld_pad), kind=kind, seed=seed + 17 * i, alpha=alpha, beta=beta)
        dA, dB, dC = c.device()
        cases.append(c)
        dev.append((dA, dB, dC))
        probs.append(sa.problem(m, n, k, alpha, dA, c.lda, dB, c.ldb, beta, dC, c.ldc))
    return cases, dev, probs


@pytest.mark.parametrize("form", [sa.FORM_F32, sa.FORM_BF16X3])
def test_grouped_matches_single_launch_bit_exact_and_oracle(form):
    import torch
    shapes = _log_uniform_shapes(64, seed=11 + form) + [(1, 1, 1), (33, 1, 65), (1, 300, 7)]
    cases, dev, probs = _make_group(shapes, seed=100, alpha=1.5, beta=0.5, ld_pad=3)
    configs = [sa.autotune_lookup(form, c.m, c.n, c.k, grouped=True)[0] for c in cases]
    log(f"grouped form {form}: {len(probs)} problems")
    plan = sa.GroupedPlan(len(probs), form)
    plan.set(probs)
    plan.run()
    torch.cuda.synchronize()
    tol = orc.fp32_tol_mult if form == sa.FORM_F32 else orc.bf16x3_tol_mult
    for c, (dA, dB, dC), cfg in zip(cases, dev, configs):
        got = download(dC)
        single = upload(c.C0)
        sa.run_config(form, cfg, c.m, c.n, c.k, c.alpha, dA, c.lda, dB, c.ldb, c.beta, single,
                      c.ldc)
        assert np.array_equal(got, download(single)), (c.m, c.n, c.k)
        st = check(c, got, tol(c.k))
        assert st["n_fail"] == 0 and st["padding_ok"], ((c.m, c.n, c.k), st)
    plan.destroy()


def test_graph_replay_equals_direct_and_tracks_set():
    """Capture once, replay several times with eager work between replays;
    every replay equals the eager launch.  A later set() with other problems
    reuses the same graph."""
    import torch
    stream = torch.cuda.Stream()
    shapes = _log_uniform_shapes(40, seed=5)
    cases, dev, probs = _make_group(shapes, seed=300)
    plan = sa.GroupedPlan(64, sa.FORM_F32)
    with torch.cuda.stream(stream):
        plan.set(probs, stream)
        plan.run(stream)
        stream.synchronize()
        eager = [download(d[2]).copy() for d in dev]
        plan.capture(stream)
        for rep in range(3):
            for d in dev:
                d[2].fill_(-7.0)  # eager work between replays
            stream.synchronize()
            plan.replay(stream)
            stream.synchronize()
            for e, d in zip(eager, dev):
                assert np.array_equal(download(d[2]), e), rep
        # re-fill the plan with a different set of problems; the graph follows
        shapes2 = _log_uniform_shapes(25, seed=6)
        cases2, dev2, probs2 = _make_group(shapes2, seed=400)
        plan.set(probs2, stream)
        plan.replay(stream)
        stream.synchronize()
    for c, d in zip(cases2, dev2):
        st = check(c, download(d[2]), orc.fp32_tol_mult(c.k))
        assert st["n_fail"] == 0, ((c.m, c.n, c.k), st)
    plan.destroy()


@pytest.mark.parametrize("form", [sa.FORM_F32, sa.FORM_BF16X3])
def test_tile_queue_starts_from_zero_on_every_launch(form):
    """The persistent workgroups take tiles from a device-side counter that
    the last workgroup out resets.  Launches with fewer tiles than workgroups,
    an empty group replayed from a graph, and back-to-back launches with no
    host synchronisation between them must each see the queue at zero: a
    launch that started from a stale index would skip tiles."""
    import torch
    tol = orc.fp32_tol_mult if form == sa.FORM_F32 else orc.bf16x3_tol_mult
    stream = torch.cuda.Stream()
    tiny_cases, tiny_dev, tiny = _make_group([(20, 30, 40)], seed=700)
    big_cases, big_dev, big = _make_group(_log_uniform_shapes(120, hi=700, seed=8), seed=800)
    plan = sa.GroupedPlan(128, form)
    with torch.cuda.stream(stream):
        plan.set(tiny, stream)
        plan.run(stream)  # one tile, every other workgroup leaves at once
        plan.set(big, stream)
        for _ in range(4):  # back to back on the stream, no host sync between
            plan.run(stream)
        stream.synchronize()
        for c, d in zip(big_cases + tiny_cases, big_dev + tiny_dev):
            st = check(c, download(d[2]), tol(c.k))
            assert st["n_fail"] == 0, ((c.m, c.n, c.k), st)
        plan.capture(stream)
        plan.set([], stream)
        plan.replay(stream)  # zero tiles: every workgroup fetches once and leaves
        for d in big_dev:
            d[2].fill_(-3.0)
        plan.set(big, stream)
        plan.replay(stream)
        plan.replay(stream)
        stream.synchronize()
    for c, d in zip(big_cases, big_dev):
        st = check(c, download(d[2]), tol(c.k))
        assert st["n_fail"] == 0, ((c.m, c.n, c.k), st)
    plan.destroy()


def test_one_shot_grouped_and_empty_members():
    shapes = [(100, 90, 80), (0, 5, 5), (7, 0, 3), (64, 64, 0), (129, 33, 257)]
    cases, dev, probs = _make_group(shapes, seed=500, alpha=1.0, beta=0.5)
    sa.grouped(probs, sa.FORM_F32)
    for c, d in zip(cases, dev):
        if c.m == 0 or c.n == 0:
            continue
        st = check(c, download(d[2]), orc.fp32_tol_mult(max(c.k, 1)))
        assert st["n_fail"] == 0, ((c.m, c.n, c.k), st)


def test_grouped_rejects_bad_problem_and_capacity():
    C = upload(np.zeros(16, np.float32))
    plan = sa.GroupedPlan(2, sa.FORM_F32)
    bad = [sa.problem(4, 4, 4, 1.0, C, 3, C, 4, 0.0, C, 4)]  # lda < m
    with pytest.raises(sa.SgemmError) as e:
        plan.set(bad)
    assert e.value.status == 1
    misaligned = [sa.problem(4, 4, 4, 1.0, C.data_ptr() + 2, 4, C, 4, 0.0, C, 4)]
    with pytest.raises(sa.SgemmError) as e:
        plan.set(misaligned)  # not 4-byte aligned: no float pointer is
    assert e.value.status == 1
    ok = sa.problem(4, 4, 4, 1.0, C, 4, C, 4, 0.0, C, 4)
    with pytest.raises(sa.SgemmError):
        plan.set([ok, ok, ok])  # over capacity
    with pytest.raises(sa.SgemmError) as e:
        plan.set([ok], configs=[3])  # config not carried by the grouped kernel
    assert e.value.status == 2
    fresh = sa.GroupedPlan(1, sa.FORM_F32)
    with pytest.raises(sa.SgemmError) as e:
        fresh.replay()
    assert e.value.status == 9
    plan.destroy()
    fresh.destroy()


def test_autotune_bucket_table_roundtrip():
    sa.autotune_clear()
    cfg, tf = sa.autotune_bucket(sa.FORM_F32, 512, 512, 512, iters=2)
    assert 0 <= cfg < len(sa.configs(sa.FORM_F32)) and tf > 0
    assert sa.autotune_lookup(sa.FORM_F32, 500, 400, 300) == (cfg, True)  # same bucket
    cfg_g, _ = sa.autotune_bucket(sa.FORM_BF16X3, 200, 300, 1000, iters=2, grouped=True)
    assert cfg_g in (1, 8)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "table.json")
        sa.autotune_save(path)
        sa.autotune_clear()
        assert sa.autotune_size() == 0
        sa.autotune_load(path)
        assert sa.autotune_lookup(sa.FORM_F32, 512, 512, 512) == (cfg, True)
        assert sa.autotune_lookup(sa.FORM_BF16X3, 256, 512, 1024, grouped=True) == (cfg_g, True)
    sa.autotune_clear()
