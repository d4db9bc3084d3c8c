"""sgemm_ex_noalloc: the no-allocation form of sgemm_ex / sgemm_bf16x3,
whose split-K, tail-split and stream-K partials and counters live in one
static device array per GPU.  It must give the same bits as the entry point
it mirrors (same config, same kernels, same summation order) on consecutive
calls -- the second one smaller, reusing the array -- keep the split under
graph capture (where the allocating entry point runs unsplit on a stream
without a workspace), and reject a plan that does not fit with
SGEMM_STATUS_SCRATCH_TOO_SMALL, launching nothing."""
import numpy as np
import pytest

import oracle as orc
import sgemm_amd as sa
from gpu_util import Case, download, log, upload

pytestmark = pytest.mark.gpu

SCRATCH_TOO_SMALL = 10


def _uses_scratch(form, m, n, k):
    cfg, _ = sa.autotune_lookup(form, m, n, k)
    name = sa.configs(form)[cfg]["name"]
    return any(s in name for s in ("splitk", "streamk", "tailsplit")), name


def _default(form):
    return sa.sgemm_ex if form == sa.FORM_F32 else sa.sgemm_bf16x3


def _tol(form, k):
    return orc.fp32_tol_mult(k) if form == sa.FORM_F32 else orc.bf16x3_tol_mult(k)


def _sampled_ok(c, out, form):
    rng = np.random.default_rng(c.m + c.n)
    ii = rng.integers(0, c.m, 1024).astype(np.int32)
    jj = rng.integers(0, c.n, 1024).astype(np.int32)
    ref, scale = orc.dgemm_entries(c.k, c.alpha, c.A, c.lda, c.B, c.ldb, c.beta, c.C0, c.ldc,
                                   ii, jj)
    got = out[ii + jj.astype(np.int64) * c.ldc].astype(np.float64)
    assert np.all(np.abs(got - ref) <= _tol(form, c.k) * scale)


@pytest.mark.parametrize("form", [sa.FORM_F32, sa.FORM_BF16X3])
def test_noalloc_equals_default_entry_two_calls(form):
    import torch
    torch.cuda.set_device(0)
    for i, (m, n, k) in enumerate([(1152, 1152, 2048), (512, 512, 4096)]):  # second smaller
        split, name = _uses_scratch(form, m, n, k)
        assert split, name
        c = Case(m, n, k, lda=m + 4, ldb=k, ldc=m + 8, seed=40 + i, alpha=1.5, beta=0.5)
        dA, dB = upload(c.A), upload(c.B)
        want, got = upload(c.C0), upload(c.C0)
        log(f"noalloc form {form} call {i} {(m, n, k)} config {name}")
        _default(form)(m, n, k, c.alpha, dA, c.lda, dB, c.ldb, c.beta, want, c.ldc)
        sa.sgemm_ex_noalloc(form, m, n, k, c.alpha, dA, c.lda, dB, c.ldb, c.beta, got, c.ldc)
        g = download(got)
        assert np.array_equal(g, download(want)), (form, (m, n, k))
        _sampled_ok(c, g, form)


@pytest.mark.parametrize("form", [sa.FORM_F32, sa.FORM_BF16X3])
def test_noalloc_keeps_the_split_under_capture(form):
    """Captured on a stream that never ran a split launch: the graph keeps the
    split (the allocating entry point would run the problem unsplit there)
    and every replay equals the eager call."""
    import torch
    m, n, k = 700, 900, 4096
    assert _uses_scratch(form, m, n, k)[0]
    c = Case(m, n, k, seed=77, alpha=1.0, beta=0.5)
    dA, dB = upload(c.A), upload(c.B)
    eager = upload(c.C0)
    sa.sgemm_ex_noalloc(form, m, n, k, 1.0, dA, m, dB, k, 0.5, eager, m)
    want = download(eager)
    dC = upload(c.C0)
    C0 = upload(c.C0)
    side = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    log(f"noalloc capture form {form}")
    with torch.cuda.graph(g, stream=side):
        sa.sgemm_ex_noalloc(form, m, n, k, 1.0, dA, m, dB, k, 0.5, dC, m, side)
    for rep in range(3):
        dC.copy_(C0)             # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(download(dC), want), rep
    del g


def test_noalloc_rejects_a_plan_larger_than_the_array():
    """BF16x3 stream-K at 3072^3 spreads 144 tiles of 96 k-tiles over 256
    workgroups, 3 pieces of 256 KiB each: 192 MiB of partials, more than the
    64 MiB array.  Nothing is launched: C keeps its bits."""
    m = n = k = 3072
    split, name = _uses_scratch(sa.FORM_BF16X3, m, n, k)
    assert "streamk" in name
    c = Case(m, n, k, seed=5)
    dA, dB, dC = c.device()
    st = sa.sgemm_ex_noalloc(sa.FORM_BF16X3, m, n, k, 1.0, dA, m, dB, k, 0.5, dC, m,
                             check=False)
    assert st == SCRATCH_TOO_SMALL
    assert np.array_equal(download(dC), c.C0)
    assert "scratch" in sa.status_string(st)


def test_noalloc_argument_rules_match_sgemm_ex():
    C = upload(np.ones(64, np.float32))
    assert sa.sgemm_ex_noalloc(sa.FORM_F32, 8, 8, 8, 1.0, C, 7, C, 8, 0.0, C, 8,
                               check=False) == 1          # lda < m
    assert sa.sgemm_ex_noalloc(5, 8, 8, 8, 1.0, C, 8, C, 8, 0.0, C, 8, check=False) == 2
    # k == 0: C := beta * C, no scratch involved
    assert sa.sgemm_ex_noalloc(sa.FORM_BF16X3, 8, 8, 0, 1.0, C, 8, C, 1, 0.5, C, 8) == 0
    assert np.array_equal(download(C), np.full(64, 0.5, np.float32))
