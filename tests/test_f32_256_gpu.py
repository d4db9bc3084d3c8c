"""Plain FP32 config 15: the 256x256x32 tile of four 128x128 waves (one per
SIMD), direct-to-LDS loads, operands double-buffered in registers.  Its
lanes supply the same k sets to each MFMA step, in the same order, as config
5's tile, so the two give the same bits; each case here compares them and
checks config 15 against the FP64 oracle.  Cases cover whole and partial
256x256 tiles, a single pair of k-tiles (k = 64), odd ldc, beta = 0 over a
NaN-filled C, and, where the tile's preconditions do not hold (k not a
multiple of 32, k = 32, unaligned columns), the config-9 path it runs
instead."""
import numpy as np
import pytest

import oracle as orc
import sgemm_amd as sa
from gpu_util import Case, assert_ok, check, compact, download, log, upload

pytestmark = pytest.mark.gpu


def _aligned(v):
    return (v + 3) // 4 * 4


def _run(config, c):
    dA, dB, dC = upload(c.A), upload(c.B), upload(c.C0)
    log(f"f32_256 config={config} m={c.m} n={c.n} k={c.k} lda={c.lda} ldb={c.ldb} "
        f"ldc={c.ldc} beta={c.beta}")
    st = sa.run_config(sa.FORM_F32, config, c.m, c.n, c.k, c.alpha, dA, c.lda, dB, c.ldb,
                       c.beta, dC, c.ldc, None)
    assert st == 0, st
    return download(dC)


SUPPORTED = [(256, 256, 64), (512, 768, 96), (700, 300, 128), (257, 513, 2048),
             (1000, 1000, 1024), (256, 256, 4096), (300, 1030, 160), (64, 64, 64),
             (1, 300, 256), (300, 1, 256)]


@pytest.mark.parametrize("shape", SUPPORTED)
def test_config15_matches_config5_bits(shape):
    m, n, k = shape
    c = Case(m, n, k, lda=_aligned(m), ldb=_aligned(k), ldc=m + 1, seed=m + 3 * n + k,
             alpha=1.5, beta=0.5)
    out15 = _run(15, c)
    assert_ok(check(c, out15, orc.fp32_tol_mult(k)), f"config 15 {shape}")
    out5 = _run(5, c)
    assert np.array_equal(compact(out15, m, n, c.ldc), compact(out5, m, n, c.ldc)), shape
    # entries of C outside the m x n result (ldc > m) are left alone
    assert np.array_equal(out15.reshape(-1), out5.reshape(-1)), shape


def test_config15_beta_zero_does_not_read_nan_c():
    for (m, n, k) in [(512, 256, 128), (300, 260, 64)]:
        c = Case(m, n, k, lda=_aligned(m), ldb=_aligned(k), seed=9, alpha=1.25, beta=0.0,
                 c_fill=np.nan)
        out = _run(15, c)
        assert not np.isnan(compact(out, m, n, c.ldc)).any()
        assert_ok(check(c, out, orc.fp32_tol_mult(k)), f"beta=0 config 15 {(m, n, k)}")


def test_config15_integer_kat_bit_exact():
    """Integer inputs in [-4, 4]: every product and partial sum is exact."""
    for (m, n, k) in [(512, 512, 512), (260, 700, 96)]:
        c = Case(m, n, k, lda=_aligned(m), ldb=_aligned(k), ldc=m + 3, kind="int", seed=m + k)
        out = _run(15, c)
        ref, _ = c.reference()
        assert np.array_equal(compact(out, m, n, c.ldc).astype(np.float64), ref), (m, n, k)


@pytest.mark.parametrize("shape,lda_pad", [((300, 200, 100), 0), ((256, 256, 32), 0),
                                           ((512, 512, 512), 1)])
def test_config15_outside_preconditions_runs_config9(shape, lda_pad):
    m, n, k = shape
    lda = _aligned(m) + lda_pad
    c = Case(m, n, k, lda=lda, ldb=_aligned(k), ldc=m, seed=m + k, alpha=1.5, beta=0.5)
    out15 = _run(15, c)
    assert_ok(check(c, out15, orc.fp32_tol_mult(k)), f"config 15 fallback {shape}")
    out9 = _run(9, c)
    assert np.array_equal(out15, out9), shape


def test_config15_at_scale_matches_config5_and_oracle():
    """4096 x 4096 x 2048 (one full round of 256 tiles) and a grid with a
    partial row and column of tiles: bits equal to config 5, and config 15
    within the FP32 bound of the FP64 oracle on sampled entries."""
    for (m, n, k) in [(4096, 4096, 2048), (3000, 2100, 1024)]:
        c = Case(m, n, k, seed=k + 7, alpha=1.5, beta=0.5)
        out15 = _run(15, c)
        out5 = _run(5, c)
        assert np.array_equal(out15, out5), (m, n, k)
        rng = np.random.default_rng(m)
        idx = rng.integers(0, m * n, size=4096)
        rows, cols = idx % m, idx // m
        ref, scale = orc.dgemm_entries(k, c.alpha, c.A, c.lda, c.B, c.ldb, c.beta, c.C0, c.ldc,
                                       rows, cols)
        got = out15[cols * c.ldc + rows].astype(np.float64)
        err = np.abs(got - ref) / np.maximum(scale, np.finfo(np.float64).tiny)
        assert err.max() <= orc.fp32_tol_mult(k), err.max()


def test_default_entry_point_picks_config15_for_whole_rounds():
    """The heuristic takes config 15 where its tiles fill whole rounds of the
    CUs (aligned columns, k a multiple of 32), config 9 where they would not."""
    base = 1 << 20  # any 16-byte aligned address: the query reads no memory
    for (m, n, k) in [(8192, 8192, 8192), (16384, 16384, 16384), (4096, 4096, 4096)]:
        assert sa.launch_config(sa.FORM_F32, m, n, k, base, m, base, k) == 15, (m, n, k)
    assert sa.launch_config(sa.FORM_F32, 8192, 8192, 8192, base, 8193, base, 8192) == 9
    assert sa.launch_config(sa.FORM_F32, 8192, 8192, 8200, base, 8192, base, 8200) == 9


# ----------------------------------------------- config 16: stream-K tail
def _sampled(c, out, n_samples, seed):
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, c.m * c.n, size=n_samples)
    rows, cols = idx % c.m, idx // c.m
    ref, scale = orc.dgemm_entries(c.k, c.alpha, c.A, c.lda, c.B, c.ldb, c.beta, c.C0, c.ldc,
                                   rows, cols)
    got = out[cols * c.ldc + rows].astype(np.float64)
    err = np.abs(got - ref) / np.maximum(scale, np.finfo(np.float64).tiny)
    assert err.max() <= orc.fp32_tol_mult(c.k), err.max()


@pytest.mark.parametrize("m,n,k", [
    (1000, 1100, 2048),   # 20 tiles of 256x256, no full round: every tile cut
    (2816, 2816, 1536),   # 121 tiles
    (4352, 4352, 1024),   # 289 tiles: one round and 33
    (300, 260, 96),       # one tile of three k-tiles
])
def test_config16_deterministic_and_in_tolerance(m, n, k):
    """Stream-K over the tiles of a partial round (or all tiles): within the
    FP32 bound of the FP64 oracle, and bit-identical on repeat with another
    problem of the same shape in between (partial slots and counters are
    reused)."""
    c = Case(m, n, k, ldc=m + 1, seed=3 * m + n + k, alpha=1.25, beta=0.5)
    first = _run(16, c)
    if m * n <= (1 << 21):
        assert_ok(check(c, first, orc.fp32_tol_mult(k)), f"config 16 {(m, n, k)}")
    else:
        _sampled(c, first, 4096, k)
    other = Case(m, n, k, ldc=m + 1, seed=m + n + k + 13, alpha=1.0, beta=0.0)
    _sampled(other, _run(16, other), 2048, k + 5)
    again = _run(16, c)
    assert np.array_equal(first, again), (m, n, k)


def test_config16_integer_kat_equals_whole_tiles():
    """Integer inputs: every partial sum is an exact integer, so the stream-K
    tail equals config 5's whole tiles bit for bit."""
    for (m, n, k) in [(2816, 2432, 1536), (4352, 4352, 1024)]:
        c = Case(m, n, k, kind="int", seed=29, alpha=1.5, beta=0.5)
        assert np.array_equal(_run(5, c), _run(16, c)), (m, n, k)


def test_config16_graph_replays_and_capture_fallback():
    """Captured after a warm-up call on the stream, config 16 replays
    bit-identically to its eager result; captured on a fresh stream (no
    workspace yet) it runs config 15's whole tiles, config 5's bits."""
    import torch
    m, n, k = 2816, 2560, 1024
    c = Case(m, n, k, seed=81, alpha=1.0, beta=0.0)
    dA, dB, dC = c.device()
    log(f"config 16 graph m={m} n={n} k={k}")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        sa.run_config(0, 16, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, s)
        s.synchronize()
        eager = download(dC).copy()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            sa.run_config(0, 16, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m,
                          torch.cuda.current_stream())
    for rep in range(3):
        dC.fill_(5.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(download(dC), eager), rep
    sa.run_config(0, 5, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m)
    torch.cuda.synchronize()
    whole = download(dC).copy()
    assert not np.array_equal(whole, eager)  # the tiles were cut
    fresh = torch.cuda.Stream()
    g2 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g2, stream=fresh):
        sa.run_config(0, 16, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m,
                      torch.cuda.current_stream())
    for rep in range(2):
        dC.fill_(-2.0)
        torch.cuda.synchronize()
        g2.replay()
        torch.cuda.synchronize()
        assert np.array_equal(download(dC), whole), rep


def test_config16_concurrent_streams_match_alone():
    """Each stream has its own workspace: two config-16 launches of different
    problems on two streams at once give the bits each gives alone."""
    import torch
    shapes = [(2816, 2560, 1536), (2432, 2600, 1024)]
    cases = [Case(m, n, k, seed=95 + i, alpha=1.0, beta=0.0) for i, (m, n, k) in enumerate(shapes)]
    devs = [c.device() for c in cases]
    alone = []
    for c, (dA, dB, dC) in zip(cases, devs):
        log(f"config 16 alone m={c.m} n={c.n} k={c.k}")
        sa.run_config(0, 16, c.m, c.n, c.k, 1.0, dA, c.m, dB, c.k, 0.0, dC, c.m)
        torch.cuda.synchronize()
        alone.append(download(dC).copy())
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for rep in range(3):
        for (_, _, dC) in devs:
            dC.fill_(3.0)
        torch.cuda.synchronize()
        for c, (dA, dB, dC), s in zip(cases, devs, streams):
            sa.run_config(0, 16, c.m, c.n, c.k, 1.0, dA, c.m, dB, c.k, 0.0, dC, c.m, s)
        torch.cuda.synchronize()
        for (_, _, dC), want in zip(devs, alone):
            assert np.array_equal(download(dC), want), rep


def test_config16_noalloc_runs_whole_tiles():
    """The no-allocation entry point never takes the stream-K workspace:
    a config-16 pick there runs config 15's whole tiles."""
    m, n, k = 4352, 4352, 1024
    c = Case(m, n, k, seed=7, alpha=1.5, beta=0.5)
    dA, dB, dC = c.device()
    sa.autotune_set(sa.FORM_F32, m, n, k, 16)
    try:
        log(f"config 16 noalloc m={m} n={n} k={k}")
        sa.sgemm_ex_noalloc(sa.FORM_F32, m, n, k, 1.5, dA, m, dB, k, 0.5, dC, m)
        got = download(dC)
    finally:
        sa.autotune_clear()
    assert np.array_equal(got, _run(15, c))
