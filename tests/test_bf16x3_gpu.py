"""GPU parity of the BF16x3 split-precision form against the FP64 oracle.

Tolerance: the plain-FP32 bound widened by the terms the split drops,
    |C - C_ref|_ij <= 2 * (gamma_{k+2} + 3 * 2^-16) * scale_ij,
scale = |alpha| (|A||B|) + |beta| |C0|.  (|lo*lo| <= 2^-16|ab| and the two
residual products <= 2^-16|ab| each.)  For k >= ~800 the added term is below
gamma_{k+2}, so at the headline sizes the form meets the plain FP32 bound
itself -- test_error_within_fp32_bound checks exactly that.
Integer-valued inputs are BF16-exact (lo = 0), so those cases are bit-exact.
"""
import json
import os

import numpy as np
import pytest

import oracle as orc
import sgemm_amd as sa
from gpu_util import (Case, assert_ok, check, compact, download, log, padding_untouched, run_form,
                      upload)

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gemm_cases.json")
N_X3_CONFIGS = 14
X3_PACKED_CONFIG = 12
X3_PACKED_STREAMK_CONFIG = 13


def test_config_table():
    assert len(sa.configs(sa.FORM_BF16X3)) == N_X3_CONFIGS


@pytest.mark.parametrize("config", range(N_X3_CONFIGS))
def test_golden_cases(config):
    with open(GOLDEN) as f:
        golden = json.load(f)["cases"]
    for g in golden:
        case = Case(g["m"], g["n"], g["k"], g["lda"], g["ldb"], g["ldc"], alpha=g["alpha"],
                    beta=g["beta"])
        case.A = np.array(g["A"], np.float32)
        case.B = np.array(g["B"], np.float32)
        case.C0 = np.array(g["C0"], np.float32)
        out = run_form(case, sa.FORM_BF16X3, config)
        got = compact(out, g["m"], g["n"], g["ldc"])
        if g["kind"] == "int":
            assert np.array_equal(got, np.array(g["expected_f32_rounded"], np.float32)), g["name"]
        assert_ok(check(case, out, orc.bf16x3_tol_mult(g["k"])), f"{g['name']} config {config}")


def test_split_matches_rne_definition():
    """K = 1 and B = 1: C = hi(a) + lo(a) exactly, which pins both GPU
    conversions to the CPU round-to-nearest-even definition."""
    m = 4096
    a = orc.fill("wide", m, 77)
    # add exact ties and near-ties of the BF16 grid
    specials = np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(1.0 + 3 * 2.0 ** -8),
                         1.0 + 2.0 ** -8 + 2.0 ** -20, 3.0e-39, 1.0e-30, 65504.0, 0.0],
                        dtype=np.float32)
    a[: specials.size] = specials
    hi, lo = orc.bf16x3_split(a)
    want = np.array([orc.bf16_to_f32(int(h)) + orc.bf16_to_f32(int(l)) for h, l in zip(hi, lo)],
                    dtype=np.float32)
    dA, dB, dC = upload(a), upload(np.ones(1, np.float32)), upload(np.zeros(m, np.float32))
    for config in range(N_X3_CONFIGS):
        log(f"bf16x3 split probe config {config}")
        sa.run_config(sa.FORM_BF16X3, config, m, 1, 1, 1.0, dA, m, dB, 1, 0.0, dC, m)
        assert np.array_equal(download(dC), want), config


@pytest.mark.parametrize("config", range(N_X3_CONFIGS))
def test_integer_kat_bit_exact(config):
    for (m, n, k) in [(300, 257, 129), (129, 130, 1000), (64, 64, 8192)]:
        c = Case(m, n, k, lda=m + 3, ldb=k + 1, ldc=m + 5, kind="int", seed=m + n + k)
        out = run_form(c, sa.FORM_BF16X3, config)
        ref, _ = c.reference()
        assert np.array_equal(compact(out, m, n, c.ldc).astype(np.float64), ref), (m, n, k)


@pytest.mark.parametrize("config", range(N_X3_CONFIGS))
def test_edge_sweep(config):
    rng = np.random.default_rng(99)
    shapes = [(s, s, s) for s in range(1, 40, 3)]
    shapes += [tuple(int(v) for v in rng.integers(1, 200, size=3)) for _ in range(40)]
    shapes += [(127, 129, 31), (128, 128, 32), (129, 127, 33), (65, 63, 65)]
    for (m, n, k) in shapes:
        c = Case(m, n, k, lda=m + 1, ldb=k + 2, ldc=m + 3, seed=m + 5 * n + 7 * k)
        out = run_form(c, sa.FORM_BF16X3, config)
        assert_ok(check(c, out, orc.bf16x3_tol_mult(k)), f"edge {(m, n, k)} config {config}")


def test_beta_zero_nan_c_and_k_zero():
    c = Case(150, 70, 64, seed=4, beta=0.0, c_fill=np.nan)
    out = run_form(c, sa.FORM_BF16X3)
    assert not np.isnan(compact(out, c.m, c.n, c.ldc)).any()
    assert_ok(check(c, out, orc.bf16x3_tol_mult(c.k)), "beta=0 NaN C")
    c = Case(40, 30, 0, seed=4, alpha=2.0, beta=0.5)
    out = run_form(c, sa.FORM_BF16X3)
    assert np.array_equal(compact(out, 40, 30, c.ldc), np.float32(0.5) * compact(c.C0, 40, 30, c.ldc))


@pytest.mark.parametrize("config,m,n,k", [(6, 190, 77, 4096), (7, 190, 77, 4096),
                                           (10, 190, 77, 4096), (9, 4352, 4096, 1024),
                                           (11, 3000, 2000, 1024), (11, 4352, 4096, 1024)])
def test_beta_zero_nan_c_with_split_k(config, m, n, k):
    """beta == 0 never reads C on the split paths (split-K and the tail
    split's reduction run the same epilogue)."""
    c = Case(m, n, k, seed=7, beta=0.0, c_fill=np.nan)
    out = run_form(c, sa.FORM_BF16X3, config)
    assert not np.isnan(compact(out, m, n, c.ldc)).any()
    _sampled_check_x3(c, out, n_samples=2048, seed=config)


@pytest.mark.parametrize("n", [1024, 2048])
def test_wide_range_error_report(n):
    """Config-3 style inputs (sign * 2^u, u in [-8, 8)): both forms against FP64.
    Prints max/mean componentwise and relative errors of each form."""
    res = {}
    c = Case(n, n, n, kind="wide", seed=3000 + n, alpha=1.0, beta=0.0)
    ref = c.reference()
    for form, tol in ((sa.FORM_F32, orc.fp32_tol_mult(n)), (sa.FORM_BF16X3, orc.bf16x3_tol_mult(n))):
        out = run_form(c, form)
        st = check(c, out, tol, ref=ref)
        assert_ok(st, f"wide {n} form {form}")
        res[form] = st
    log(f"wide-range {n}^3: f32 max/mean comp {res[0]['max_err']:.3e}/{res[0]['mean_err']:.3e} "
        f"max/mean rel {res[0]['max_rel']:.3e}/{res[0]['mean_rel']:.3e} | bf16x3 max/mean comp "
        f"{res[1]['max_err']:.3e}/{res[1]['mean_err']:.3e} max/mean rel "
        f"{res[1]['max_rel']:.3e}/{res[1]['mean_rel']:.3e}")


@pytest.mark.parametrize("n", [8192, 16384])
def test_error_against_plain_fp32_same_entries(n):
    """Both forms on the same wide-range inputs (sign * 2^u, u in [-8, 8)) at
    the split-precision benchmark's sizes, against the host FP64 oracle on
    the same 8192 sampled entries.

    The benchmark's target -- BF16x3 error no larger than the plain form's --
    is NOT met on these entries, and this test states by how much: the
    two-term split drops a residual of ~2^-17 |x| per operand, which puts the
    BF16x3 mean componentwise error at 2-3.5x the plain kernel's and its
    sampled max at about 1.1-1.4x (measured on MI355X: 8192^3 max 1.06e-6 vs
    0.84e-6, mean 2.0e-7 vs 6.6e-8; 16384^3 max 6.8e-7 vs 5.7e-7, mean 1.5e-7
    vs 6.5e-8).  Asserted: both forms inside the plain FP32 stated bound
    2 gamma_{k+2}; BF16x3's sampled max within 2x and its mean within 4x of
    the plain form's on the same entries."""
    c = Case(n, n, n, kind="wide", seed=3 * n + 1, alpha=1.0, beta=0.0)
    rng = np.random.default_rng(n)
    ii = rng.integers(0, n, 8192).astype(np.int32)
    jj = rng.integers(0, n, 8192).astype(np.int32)
    ref, scale = orc.dgemm_entries(n, 1.0, c.A, n, c.B, n, 0.0, None, n, ii, jj)
    dA, dB = upload(c.A), upload(c.B)
    err = {}
    for form, fn in ((sa.FORM_F32, sa.sgemm_ex), (sa.FORM_BF16X3, sa.sgemm_bf16x3)):
        dC = upload(np.zeros(n * n, np.float32))
        log(f"error comparison {n}^3 form {form}")
        fn(n, n, n, 1.0, dA, n, dB, n, 0.0, dC, n)
        got = download(dC)[ii + jj.astype(np.int64) * n].astype(np.float64)
        del dC
        e = np.abs(got - ref) / scale
        assert np.all(np.isfinite(e)) and np.all(e <= orc.fp32_tol_mult(n)), (form, e.max())
        err[form] = (float(e.max()), float(e.mean()))
    (fmax, fmean), (xmax, xmean) = err[sa.FORM_F32], err[sa.FORM_BF16X3]
    log(f"{n}^3 wide, 8192 sampled entries: plain max {fmax:.3e} mean {fmean:.3e} | "
        f"bf16x3 max {xmax:.3e} mean {xmean:.3e} | ratio max {xmax / fmax:.2f} "
        f"mean {xmean / fmean:.2f}")
    assert xmax <= 2.0 * fmax and xmean <= 4.0 * fmean, err


def test_error_within_fp32_bound_at_8192():
    """At k = 8192 the BF16x3 result must meet the plain FP32 stated bound
    2 gamma_{k+2} scale (sampled entries against FP64)."""
    n = 8192
    c = Case(n, n, n, kind="wide", seed=8192, alpha=1.0, beta=0.0)
    out = run_form(c, sa.FORM_BF16X3)
    rng = np.random.default_rng(1)
    ii = rng.integers(0, n, 4096).astype(np.int32)
    jj = rng.integers(0, n, 4096).astype(np.int32)
    ref, scale = orc.dgemm_entries(n, 1.0, c.A, n, c.B, n, 0.0, None, n, ii, jj)
    got = out[ii + jj.astype(np.int64) * n].astype(np.float64)
    e = np.abs(got - ref) / scale
    log(f"bf16x3 8192^3 wide sampled max comp err {e.max():.3e} (fp32 bound "
        f"{orc.fp32_tol_mult(n):.3e})")
    assert np.all(e <= orc.fp32_tol_mult(n))


# ------------------------------------------------------------------- split-K
@pytest.mark.parametrize("config", (6, 7, 10))
def test_splitk_deterministic_and_in_tolerance(config):
    """Split-K configs (tiles of configs 0 and 1): partials summed in split
    order, so repeated launches are bit-identical; a different problem of the
    same shape in between reuses the partial slots (a stale read would fail
    its tolerance check)."""
    for (m, n, k) in [(256, 256, 8192), (300, 200, 5000), (129, 257, 777), (1024, 64, 4096)]:
        c = Case(m, n, k, seed=m + k, alpha=1.25, beta=0.5)
        first = run_form(c, sa.FORM_BF16X3, config)
        assert_ok(check(c, first, orc.bf16x3_tol_mult(k)), f"split-K {(m, n, k)} config {config}")
        other = Case(m, n, k, kind="wide", seed=m + k + 1000, alpha=1.0, beta=0.0)
        out = run_form(other, sa.FORM_BF16X3, config)
        assert_ok(check(other, out, orc.bf16x3_tol_mult(k)),
                  f"split-K 2nd {(m, n, k)} config {config}")
        again = run_form(c, sa.FORM_BF16X3, config)
        assert np.array_equal(first, again), (m, n, k)


def test_splitk_integer_kat_bit_exact():
    """Integer inputs are BF16-exact and every partial sum is an exact
    integer, so the split result equals the unsplit one bit for bit."""
    m, n, k = 96, 80, 4096
    c = Case(m, n, k, kind="int", seed=9, alpha=1.5, beta=0.5)
    a = run_form(c, sa.FORM_BF16X3, 0)
    b = run_form(c, sa.FORM_BF16X3, 6)
    assert np.array_equal(a, b)
    assert_ok(check(c, b, orc.bf16x3_tol_mult(k)), "split-K integer KAT")
    # config 10 splits config 8's tile (16 k-tiles per split at least)
    c = Case(256, 384, 4096, kind="int", seed=10, alpha=1.5, beta=0.5)
    assert np.array_equal(run_form(c, sa.FORM_BF16X3, 8), run_form(c, sa.FORM_BF16X3, 10))


# ------------------------------------------------------------ tail split
X3_TAIL_CONFIG = 9  # config 2's 256x256 tile; the tiles of a partial last round split k


@pytest.mark.parametrize("m,n,k,ld_pad", [
    (4100, 3900, 1000, 3),   # 17x16 = 272 tiles of 256x256: 16 past one 256-CU round
    (7168, 7168, 2048, 0),   # 784 tiles: 16 past three rounds
])
def test_tail_split_deterministic_and_in_tolerance(m, n, k, ld_pad):
    """In the BF16x3 tolerance against FP64 (every entry below 2^24 of them,
    sampled above), bit-identical on repeat with another problem of the
    same shape in between (the partial slots are reused)."""
    c = Case(m, n, k, lda=m + ld_pad, ldb=k + ld_pad, ldc=m + ld_pad, seed=m + n + k,
             alpha=1.25, beta=0.5)
    first = run_form(c, sa.FORM_BF16X3, X3_TAIL_CONFIG)
    if m * n <= (1 << 24):
        assert_ok(check(c, first, orc.bf16x3_tol_mult(k)), f"tail split {(m, n, k)}")
    else:
        _sampled_check_x3(c, first, n_samples=4096, seed=k)
    assert padding_untouched(first, c.C0, m, n, c.ldc)
    other = Case(m, n, k, lda=m + ld_pad, ldb=k + ld_pad, ldc=m + ld_pad, kind="wide",
                 seed=m + n + k + 7, alpha=1.0, beta=0.0)
    out = run_form(other, sa.FORM_BF16X3, X3_TAIL_CONFIG)
    _sampled_check_x3(other, out, n_samples=2048, seed=k + 1)
    again = run_form(c, sa.FORM_BF16X3, X3_TAIL_CONFIG)
    assert np.array_equal(first, again), (m, n, k)


def _sampled_check_x3(c, out, n_samples, seed):
    rng = np.random.default_rng(seed)
    ii = np.concatenate([rng.integers(0, c.m, n_samples), [0, c.m - 1, 0, c.m - 1]]).astype(np.int32)
    jj = np.concatenate([rng.integers(0, c.n, n_samples), [0, 0, c.n - 1, c.n - 1]]).astype(np.int32)
    ref, scale = orc.dgemm_entries(c.k, c.alpha, c.A, c.lda, c.B, c.ldb, c.beta, c.C0, c.ldc, ii, jj)
    got = out[ii + jj.astype(np.int64) * c.ldc].astype(np.float64)
    err = np.abs(got - ref)
    tol = orc.bf16x3_tol_mult(c.k) * scale
    assert np.all(np.isfinite(got)) and np.all(err <= tol), float(np.max(err / tol))


def test_tail_split_integer_kat_equals_unsplit():
    """Integer inputs are BF16-exact and every partial sum is an exact
    integer, so the tail split equals config 2 (no split) bit for bit."""
    m, n, k = 4352, 4096, 2048  # 272 tiles: 16 split tiles
    c = Case(m, n, k, kind="int", seed=19, alpha=1.5, beta=0.5)
    a = run_form(c, sa.FORM_BF16X3, 2)
    b = run_form(c, sa.FORM_BF16X3, X3_TAIL_CONFIG)
    assert np.array_equal(a, b)


def test_tail_split_graph_replays_and_capture_fallback():
    """Captured after a warm-up call on the stream, the tail split replays
    bit-identically to its eager result; captured on a fresh stream (no
    workspace yet) it runs config 2's kernel unsplit, bit for bit."""
    import torch
    m, n, k = 4352, 4096, 1024
    c = Case(m, n, k, seed=78, alpha=1.0, beta=0.0)
    dA, dB, dC = c.device()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        sa.run_config(1, X3_TAIL_CONFIG, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, s)
        s.synchronize()
        eager = download(dC).copy()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            sa.run_config(1, X3_TAIL_CONFIG, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m,
                          torch.cuda.current_stream())
    for rep in range(3):
        dC.fill_(5.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(download(dC), eager), rep
    sa.run_config(1, 2, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m)
    torch.cuda.synchronize()
    unsplit = download(dC).copy()
    assert not np.array_equal(unsplit, eager)  # the tail tiles were split
    fresh = torch.cuda.Stream()
    g2 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g2, stream=fresh):
        sa.run_config(1, X3_TAIL_CONFIG, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m,
                      torch.cuda.current_stream())
    for rep in range(2):
        dC.fill_(-2.0)
        torch.cuda.synchronize()
        g2.replay()
        torch.cuda.synchronize()
        assert np.array_equal(download(dC), unsplit), rep


# ------------------------------------------------------------ stream-K
X3_STREAMK_CONFIG = 11  # config 2's 256x256 tile; stream-K over a partial round's tiles


@pytest.mark.parametrize("m,n,k,ld_pad", [
    (1000, 1100, 2000, 3),   # 20 tiles, no full round: every tile cut into many pieces
    (3000, 3100, 1501, 0),   # 156 tiles; ragged last k-tile
    (4352, 4096, 1024, 0),   # 272 tiles: 16 past one 256-CU round
    (5120, 5120, 640, 1),    # 400 tiles: 144 past one round
])
def test_streamk_deterministic_and_in_tolerance(m, n, k, ld_pad):
    """In the BF16x3 tolerance against FP64 (every entry below 2^24 of them,
    sampled above), bit-identical on repeat with another problem of the same
    shape in between (the pieces' partial slots and counters are reused)."""
    c = Case(m, n, k, lda=m + ld_pad, ldb=k + ld_pad, ldc=m + ld_pad, seed=m + 2 * n + k,
             alpha=1.25, beta=0.5)
    first = run_form(c, sa.FORM_BF16X3, X3_STREAMK_CONFIG)
    if m * n <= (1 << 21):
        assert_ok(check(c, first, orc.bf16x3_tol_mult(k)), f"stream-K {(m, n, k)}")
    else:
        _sampled_check_x3(c, first, n_samples=4096, seed=k)
    assert padding_untouched(first, c.C0, m, n, c.ldc)
    other = Case(m, n, k, lda=m + ld_pad, ldb=k + ld_pad, ldc=m + ld_pad, kind="wide",
                 seed=m + n + k + 11, alpha=1.0, beta=0.0)
    out = run_form(other, sa.FORM_BF16X3, X3_STREAMK_CONFIG)
    _sampled_check_x3(other, out, n_samples=2048, seed=k + 3)
    again = run_form(c, sa.FORM_BF16X3, X3_STREAMK_CONFIG)
    assert np.array_equal(first, again), (m, n, k)


@pytest.mark.parametrize("m,n,k", [(2560, 2304, 1536), (4352, 4096, 2048)])
def test_streamk_integer_kat_equals_unsplit(m, n, k):
    """Integer inputs are BF16-exact and every partial sum is an exact
    integer, so stream-K equals config 2 (whole tiles) bit for bit."""
    c = Case(m, n, k, kind="int", seed=23, alpha=1.5, beta=0.5)
    a = run_form(c, sa.FORM_BF16X3, 2)
    b = run_form(c, sa.FORM_BF16X3, X3_STREAMK_CONFIG)
    assert np.array_equal(a, b)


def test_streamk_graph_replays_and_capture_fallback():
    """Captured after a warm-up call on the stream, stream-K replays
    bit-identically to its eager result; captured on a fresh stream (no
    workspace yet) it runs config 2's kernel on whole tiles, bit for bit."""
    import torch
    m, n, k = 3000, 3100, 1536
    c = Case(m, n, k, seed=79, alpha=1.0, beta=0.0)
    dA, dB, dC = c.device()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        sa.run_config(1, X3_STREAMK_CONFIG, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, s)
        s.synchronize()
        eager = download(dC).copy()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            sa.run_config(1, X3_STREAMK_CONFIG, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m,
                          torch.cuda.current_stream())
    for rep in range(3):
        dC.fill_(5.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(download(dC), eager), rep
    sa.run_config(1, 2, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m)
    torch.cuda.synchronize()
    whole = download(dC).copy()
    assert not np.array_equal(whole, eager)  # the tiles were cut
    fresh = torch.cuda.Stream()
    g2 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g2, stream=fresh):
        sa.run_config(1, X3_STREAMK_CONFIG, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m,
                      torch.cuda.current_stream())
    for rep in range(2):
        dC.fill_(-2.0)
        torch.cuda.synchronize()
        g2.replay()
        torch.cuda.synchronize()
        assert np.array_equal(download(dC), whole), rep


def test_streamk_concurrent_streams_match_sequential():
    """Each stream has its own stream-K workspace (partials and counters):
    two stream-K launches of different problems running at the same time on
    two streams give the bits each gives alone."""
    import torch
    shapes = [(3000, 3100, 1536), (2600, 2700, 2048)]
    cases = [Case(m, n, k, seed=90 + i, alpha=1.0, beta=0.0) for i, (m, n, k) in enumerate(shapes)]
    devs = [c.device() for c in cases]
    alone = []
    for c, (dA, dB, dC) in zip(cases, devs):
        sa.run_config(1, X3_STREAMK_CONFIG, c.m, c.n, c.k, 1.0, dA, c.m, dB, c.k, 0.0, dC, c.m)
        torch.cuda.synchronize()
        alone.append(download(dC).copy())
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for rep in range(3):
        for (_, _, dC) in devs:
            dC.fill_(3.0)
        torch.cuda.synchronize()
        for c, (dA, dB, dC), st in zip(cases, devs, streams):
            sa.run_config(1, X3_STREAMK_CONFIG, c.m, c.n, c.k, 1.0, dA, c.m, dB, c.k, 0.0, dC,
                          c.m, st)
        torch.cuda.synchronize()
        for i, (_, _, dC) in enumerate(devs):
            assert np.array_equal(download(dC), alone[i]), (rep, i)


# ------------------------------------------------- single-set tiles agree
@pytest.mark.parametrize("m,n,k,pad", [(300, 260, 1000, 3), (513, 257, 97, 0), (256, 512, 4096, 1)])
def test_single_set_tiles_agree_bit_for_bit(m, n, k, pad):
    """Configs 2 (256x256, eight waves, each block row a region of its
    interleaved k loop) and 8 (128x128, two per CU, each block row split into
    two regions) accumulate every C element as al*bh, ah*bl, ah*bh per 32-deep
    k-tile in k order, so their results are the same bits; the shapes run the
    interleaved loop, its hand-over to the general loop and a ragged last
    k-tile."""
    c = Case(m, n, k, lda=m + pad, ldb=k + pad, ldc=m + pad, kind="wide", seed=m * 7 + k,
             alpha=1.5, beta=0.5)
    a = run_form(c, sa.FORM_BF16X3, 2)
    b = run_form(c, sa.FORM_BF16X3, 8)
    assert_ok(check(c, a, orc.bf16x3_tol_mult(k)), f"config 2 {(m, n, k)}")
    assert np.array_equal(a, b), (m, n, k)


@pytest.mark.parametrize("m,n,k,pad", [(300, 260, 1000, 3), (513, 257, 97, 0), (1, 1, 1, 0),
                                       (17, 300, 33, 2), (256, 512, 4096, 1), (640, 300, 2049, 0)])
def test_packed_matches_config2_bit_for_bit(m, n, k, pad):
    """Config 12 splits A and B once into packed hi/lo planes (rows past the
    operand and k past K zero) and runs the 256x256 tile on them; every
    accumulator takes al*bh, ah*bl, ah*bh per 32-deep k-tile in k order, as
    config 2 does, so the bits agree.  Ragged panels and k-tiles, padded
    leading dimensions, and the FP64 oracle."""
    c = Case(m, n, k, lda=m + pad, ldb=k + pad, ldc=m + pad, kind="wide", seed=m * 5 + k,
             alpha=1.5, beta=0.5)
    a = run_form(c, sa.FORM_BF16X3, 2)
    b = run_form(c, sa.FORM_BF16X3, X3_PACKED_CONFIG)
    assert_ok(check(c, b, orc.bf16x3_tol_mult(k)), f"config 12 {(m, n, k)}")
    assert np.array_equal(a, b), (m, n, k)


def test_packed_large_ragged_matches_config2():
    """Several panels and a partial last panel in both m and n, k not a
    multiple of 32, odd leading dimensions: the same bits as config 2 (which
    the oracle checks on this shape in the sweep tests)."""
    c = Case(4097, 3071, 2049, kind="wide", seed=311, alpha=1.5, beta=0.5)
    a = run_form(c, sa.FORM_BF16X3, 2)
    b = run_form(c, sa.FORM_BF16X3, X3_PACKED_CONFIG)
    assert np.array_equal(a, b)


def test_packed_non_finite_and_beta_zero_nan_c():
    """+-Inf, NaN and values beyond the BF16 range pass through the pack's
    split as through config 2's (same bits, NaN included); with beta = 0 a
    NaN-filled C is not read."""
    c = Case(300, 280, 200, kind="wide", seed=5, alpha=1.0, beta=0.0, c_fill=np.nan)
    c.A[[0, 7, 1000, 2001]] = [np.inf, -np.inf, np.nan, np.float32(3.4e38)]
    c.B[[3, 500, 777]] = [np.float32(-3.4e38), -np.inf, np.float32(1e-40)]
    a = run_form(c, sa.FORM_BF16X3, 2)
    b = run_form(c, sa.FORM_BF16X3, X3_PACKED_CONFIG)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    clean = Case(300, 280, 200, kind="wide", seed=6, alpha=1.0, beta=0.0, c_fill=np.nan)
    out = run_form(clean, sa.FORM_BF16X3, X3_PACKED_CONFIG)
    assert not np.isnan(compact(out, 300, 280, 300)).any()
    assert_ok(check(clean, out, orc.bf16x3_tol_mult(200)), "config 12 beta 0")


def test_packed_graph_replays_and_capture_fallback():
    """Captured after a warm-up call on the stream (its pack workspace then
    exists), config 12 replays bit-identically to its eager result, with
    eager work between replays; captured on a fresh stream it allocates
    nothing and runs config 2, whose bits are the same."""
    import torch
    m, n, k = 1500, 1300, 1100
    c = Case(m, n, k, seed=83, alpha=1.0, beta=0.0)
    dA, dB, dC = c.device()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        sa.run_config(1, X3_PACKED_CONFIG, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, s)
        s.synchronize()
        eager = download(dC).copy()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            sa.run_config(1, X3_PACKED_CONFIG, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m,
                          torch.cuda.current_stream())
    other = Case(700, 600, 500, seed=84)
    oA, oB, oC = other.device()
    for rep in range(3):
        dC.fill_(5.0)
        sa.run_config(1, X3_PACKED_CONFIG, 700, 600, 500, 1.0, oA, 700, oB, 500, 0.0, oC, 700)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(download(dC), eager), rep
    fresh = torch.cuda.Stream()
    g2 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g2, stream=fresh):
        sa.run_config(1, X3_PACKED_CONFIG, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m,
                      torch.cuda.current_stream())
    for rep in range(2):
        dC.fill_(-2.0)
        torch.cuda.synchronize()
        g2.replay()
        torch.cuda.synchronize()
        assert np.array_equal(download(dC), eager), rep
    sa.workspace_release(s)


def test_packed_concurrent_streams_match_sequential():
    """Each stream has its own packed planes: two config-12 launches of
    different problems running at once on two streams give the bits each
    gives alone."""
    import torch
    shapes = [(2000, 2100, 1536), (1800, 1700, 2048)]
    cases = [Case(m, n, k, seed=95 + i, alpha=1.0, beta=0.0) for i, (m, n, k) in enumerate(shapes)]
    devs = [c.device() for c in cases]
    alone = []
    for c, (dA, dB, dC) in zip(cases, devs):
        sa.run_config(1, X3_PACKED_CONFIG, c.m, c.n, c.k, 1.0, dA, c.m, dB, c.k, 0.0, dC, c.m)
        torch.cuda.synchronize()
        alone.append(download(dC).copy())
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for rep in range(3):
        for (dA, dB, dC) in devs:
            dC.fill_(1.0)
        torch.cuda.synchronize()
        for c, (dA, dB, dC), st in zip(cases, devs, streams):
            sa.run_config(1, X3_PACKED_CONFIG, c.m, c.n, c.k, 1.0, dA, c.m, dB, c.k, 0.0, dC,
                          c.m, st)
        torch.cuda.synchronize()
        for i, (dA, dB, dC) in enumerate(devs):
            assert np.array_equal(download(dC), alone[i]), (rep, i)
    for st in streams:
        sa.workspace_release(st)


def test_default_pick_packed_and_noalloc_same_bits():
    """At 8192 x 8192 x 1024 the default BF16x3 entry point picks config 12;
    its result equals config 2's and that of the no-allocation entry point
    (which runs config 2's tile from A and B) bit for bit."""
    import torch
    m, n, k = 8192, 8192, 1024
    c = Case(m, n, k, seed=17, alpha=1.0, beta=0.0, c_fill=0.0)
    dA, dB, dC = c.device()
    assert sa.launch_config(sa.FORM_BF16X3, m, n, k, dA, m, dB, k) == X3_PACKED_CONFIG
    log("default bf16x3 8192x8192x1024")
    sa.sgemm_bf16x3(m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m)
    default = download(dC).copy()
    sa.run_config(1, 2, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m)
    assert np.array_equal(download(dC), default)
    dC.fill_(3.0)
    sa.sgemm_ex_noalloc(sa.FORM_BF16X3, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m)
    assert np.array_equal(download(dC), default)
    sa.workspace_release(None)
    torch.cuda.synchronize()


@pytest.mark.parametrize("m,n,k", [(3000, 3100, 1536), (4600, 4500, 1000), (2600, 2700, 2049)])
def test_packed_streamk_matches_config11_bit_for_bit(m, n, k):
    """Config 13 runs the packed tile with config 11's stream-K plan: the
    tiles of the partial round are cut into k-ranges, and a cut tile's pieces
    are stored in config 2's split layout and summed in k order, so the bits
    equal config 11's; also against the FP64 oracle on sampled entries."""
    c = Case(m, n, k, kind="wide", seed=m + n + k, alpha=1.5, beta=0.5)
    a = run_form(c, sa.FORM_BF16X3, X3_STREAMK_CONFIG)
    b = run_form(c, sa.FORM_BF16X3, X3_PACKED_STREAMK_CONFIG)
    assert np.array_equal(a, b), (m, n, k)
    _sampled_check_x3(c, b, 2048, m)


def test_packed_streamk_capture_fallback_and_noalloc():
    """Captured on a fresh stream (no workspace), config 13 allocates nothing
    and runs config 11's tile from A and B (or whole tiles, if its partials
    cannot grow either), and the no-allocation entry point runs config 11
    with its static partials: every result equals an eager config-13 call
    whose plan has the same cut, or config 2's whole tiles."""
    import torch
    m, n, k = 3000, 3100, 1536
    c = Case(m, n, k, seed=88, alpha=1.0, beta=0.0)
    dA, dB, dC = c.device()
    sa.run_config(1, X3_PACKED_STREAMK_CONFIG, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m)
    eager = download(dC).copy()
    sa.run_config(1, 2, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m)
    whole = download(dC).copy()
    fresh = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=fresh):
        sa.run_config(1, X3_PACKED_STREAMK_CONFIG, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m,
                      torch.cuda.current_stream())
    for rep in range(2):
        dC.fill_(-2.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(download(dC), whole), rep
    dC.fill_(7.0)
    assert sa.sgemm_ex_noalloc(sa.FORM_BF16X3, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m) == 0
    got = download(dC)
    cfg = sa.launch_config(sa.FORM_BF16X3, m, n, k, dA, m, dB, k)
    if cfg in (X3_STREAMK_CONFIG, X3_PACKED_STREAMK_CONFIG):
        assert np.array_equal(got, eager)
