"""GPU parity of the plain FP32 path (sgemm / sgemm_ex / sgemm_rm / sgemm_basic)
against the FP64 oracle, through the C ABI.

Tolerance (plain FP32, any summation order, with or without FMA):
    |C - C_ref|_ij <= 2 * gamma_{k+2} * (|alpha| (|A||B|)_ij + |beta| |C0|_ij)
gamma_k = k u / (1 - k u), u = 2^-24.  Integer-valued inputs (|x| <= 4,
k <= 8192) make every partial sum exact, so those cases must be bit-exact.

Mirrors the reference's test.cu (sgemm vs the naive kernel, alpha=1.5,
beta=0.5, test.cu:71-83) and widens it: non-square shapes, ld > dim, beta=0
with a NaN-poisoned C, k=0, invalid arguments, unaligned operands.
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
N_F32_CONFIGS = 17


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


def test_config_table():
    cfgs = sa.configs(sa.FORM_F32)
    assert len(cfgs) == N_F32_CONFIGS
    assert all(c["bm"] % 16 == 0 and c["bn"] % 16 == 0 for c in cfgs)


@pytest.mark.parametrize("config", range(N_F32_CONFIGS))
def test_golden_cases(golden, config):
    for g in golden:
        case = Case(g["m"], g["n"], g["k"], g["lda"], g["ldb"], g["ldc"], alpha=g["alpha"],
                    beta=g["beta"])
        case.A = np.array(g["A"], np.float32)
        case.B = np.array(g["B"], np.float32)
        case.C0 = np.array(g["C0"], np.float32)
        out = run_form(case, sa.FORM_F32, config)
        got = compact(out, g["m"], g["n"], g["ldc"])
        if g["kind"] == "int":
            assert np.array_equal(got, np.array(g["expected_f32_rounded"], np.float32)), g["name"]
        st = check(case, out, orc.fp32_tol_mult(g["k"]))
        assert_ok(st, f"{g['name']} config {config}")


def test_basic_kernel_bit_exact_golden(golden):
    """sgemm_basic follows kernel_basic's exact operation order."""
    for g in golden:
        dA, dB, dC = upload(np.array(g["A"])), upload(np.array(g["B"])), upload(np.array(g["C0"]))
        log(f"sgemm_basic golden {g['name']}")
        assert sa.sgemm_basic(g["m"], g["n"], g["k"], g["alpha"], dA, g["lda"], dB, g["ldb"],
                              g["beta"], dC, g["ldc"]) == 0
        got = compact(download(dC), g["m"], g["n"], g["ldc"])
        assert np.array_equal(got, np.array(g["expected_basic_f32"], np.float32)), g["name"]


def test_basic_kernel_bit_exact_vs_restatement():
    c = Case(150, 97, 301, lda=160, ldb=311, ldc=151, seed=7)
    dA, dB, dC = c.device()
    log("sgemm_basic 150x97x301")
    assert sa.sgemm_basic(c.m, c.n, c.k, c.alpha, dA, c.lda, dB, c.ldb, c.beta, dC, c.ldc) == 0
    want = c.C0.copy()
    orc.sgemm_basic_cm(c.m, c.n, c.k, c.alpha, c.A, c.lda, c.B, c.ldb, c.beta, want, c.ldc)
    assert np.array_equal(download(dC), want)


@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (1.5, 0.5)])
def test_config1_1024(alpha, beta):
    c = Case(1024, 1024, 1024, seed=101, alpha=alpha, beta=beta)
    out = run_form(c, sa.FORM_F32)
    st = check(c, out, orc.fp32_tol_mult(c.k))
    assert_ok(st, f"1024^3 alpha={alpha} beta={beta}")


SWEEP_SQUARES = [3072, 4096, 5120, 6144, 7168, 9216, 10240, 11264, 12288, 13312, 14336, 15360]


@pytest.mark.parametrize("n", SWEEP_SQUARES)
@pytest.mark.parametrize("form", [sa.FORM_F32, sa.FORM_BF16X3])
def test_square_sweep_sampled_fp64(form, n):
    """The squares of the 1024..16384 step-1024 sweep not covered below
    (1024, 2048, 8192 and 16384 are), in both forms, with the configs the
    default entry points pick: sampled entries against FP64."""
    c = Case(n, n, n, seed=n)
    out = run_form(c, form)
    tol = orc.fp32_tol_mult(n) if form == sa.FORM_F32 else orc.bf16x3_tol_mult(n)
    _sampled_check(c, out, n_samples=2048, seed=n, tol_mult=tol)


@pytest.mark.parametrize("config", range(N_F32_CONFIGS))
def test_integer_kat_bit_exact(config):
    """Integer inputs in [-4, 4]: every partial sum is exact, so any summation
    order must give exactly the FP64 result (alpha=1.5, beta=0.5 keep it exact)."""
    for (m, n, k) in [(300, 257, 129), (129, 130, 1000), (64, 64, 8192)]:
        c = Case(m, n, k, lda=m + 3, ldb=k + 1, ldc=m + 5, kind="int", seed=m + n + k)
        out = run_form(c, sa.FORM_F32, config)
        ref, _ = c.reference()
        assert np.array_equal(compact(out, m, n, c.ldc).astype(np.float64), ref), (m, n, k)


def _edge_shapes():
    shapes = [(s, s, s) for s in range(1, 34)]
    for bm in (64, 128, 256):
        for d in (-1, 0, 1):
            shapes.append((bm + d, 65 + d, 33 + d))
            shapes.append((17, bm + d, 47))
    rng = np.random.default_rng(2024)
    for _ in range(60):
        shapes.append(tuple(int(v) for v in rng.integers(1, 161, size=3)))
    return shapes


@pytest.mark.parametrize("config", range(N_F32_CONFIGS))
def test_edge_sweep(config):
    for (m, n, k) in _edge_shapes():
        c = Case(m, n, k, lda=m + (m % 3), ldb=k + (k % 2), ldc=m + 1, seed=m * 7 + n * 3 + k)
        out = run_form(c, sa.FORM_F32, config)
        st = check(c, out, orc.fp32_tol_mult(k))
        assert_ok(st, f"edge {(m, n, k)} config {config}")


def test_beta_zero_does_not_read_nan_c():
    for config in range(N_F32_CONFIGS):
        c = Case(190, 77, 45, seed=3, alpha=1.25, beta=0.0, c_fill=np.nan)
        out = run_form(c, sa.FORM_F32, config)
        got = compact(out, c.m, c.n, c.ldc)
        assert not np.isnan(got).any()
        st = check(c, out, orc.fp32_tol_mult(c.k))
        assert_ok(st, f"beta=0 NaN C config {config}")


@pytest.mark.parametrize("config,m,n,k", [(6, 190, 77, 2000), (7, 190, 77, 2000),
                                           (8, 190, 77, 2000), (10, 190, 77, 2000),
                                           (9, 3584, 3584, 1024), (11, 2000, 1900, 1536),
                                           (11, 3584, 3584, 1024)])
def test_beta_zero_nan_c_with_split_k(config, m, n, k):
    """beta == 0 never reads C on the split paths either: the last workgroup
    of a split tile runs the same epilogue (a NaN read would reach every
    entry of the tile)."""
    c = Case(m, n, k, seed=6, alpha=1.25, beta=0.0, c_fill=np.nan)
    out = run_form(c, sa.FORM_F32, config)
    assert not np.isnan(compact(out, m, n, c.ldc)).any()
    _sampled_check(c, out, n_samples=2048, seed=config)


def test_k_zero_and_alpha_zero_scale_c():
    for (k, alpha, beta) in [(0, 1.5, 0.5), (33, 0.0, 0.5), (0, 1.0, 0.0)]:
        c = Case(70, 20, k, seed=5, alpha=alpha, beta=beta)
        out = run_form(c, sa.FORM_F32)
        want = np.float32(beta) * compact(c.C0, 70, 20, c.ldc) if beta != 0 else 0.0
        assert np.array_equal(compact(out, 70, 20, c.ldc), np.broadcast_to(want, (1400,)))


def test_empty_is_noop():
    C = upload(np.arange(10, dtype=np.float32))
    for (m, n) in [(0, 5), (5, 0), (0, 0)]:
        assert sa.sgemm(m, n, 3, 1.0, C, max(1, m), C, 3, 0.0, C, max(1, m)) == 0
    assert np.array_equal(download(C), np.arange(10, dtype=np.float32))


def test_invalid_arguments_launch_nothing():
    C = upload(np.full(16, 3.0, np.float32))
    A = upload(np.ones(16, np.float32))
    bad = [
        (-1, 4, 4, 4, 4, 4), (4, -1, 4, 4, 4, 4), (4, 4, -1, 4, 4, 4),
        (4, 4, 4, 3, 4, 4), (4, 4, 4, 4, 3, 4), (4, 4, 4, 4, 4, 3),
    ]
    for (m, n, k, lda, ldb, ldc) in bad:
        assert sa.sgemm(m, n, k, 1.0, A, lda, A, ldb, 0.0, C, ldc) == 1
        assert sa.sgemm_ex(m, n, k, 1.0, A, lda, A, ldb, 0.0, C, ldc, check=False) == 1
    assert sa.sgemm(4, 4, 4, 1.0, None, 4, A, 4, 0.0, C, 4) == 1
    assert sa.sgemm(4, 4, 4, 1.0, A, 4, A, 4, 0.0, None, 4) == 1
    assert sa.run_config(0, 99, 4, 4, 4, 1.0, A, 4, A, 4, 0.0, C, 4, check=False) == 2
    assert np.array_equal(download(C), np.full(16, 3.0, np.float32))


def test_rowmajor_shim_matches_reference_semantics():
    """sgemm_rm keeps the reference's row-major meaning (src/sgemm.cuh:31-33)."""
    m, n, k = 300, 170, 90
    A = orc.fill("uniform", m * k, 11)   # row-major m x k, lda = k
    B = orc.fill("uniform", k * n, 12)   # row-major k x n, ldb = n
    C = orc.fill("uniform", m * n, 13)   # row-major m x n, ldc = n
    dA, dB, dC = upload(A), upload(B), upload(C)
    log("sgemm_rm 300x170x90")
    assert sa.sgemm_rm(m, n, k, 1.5, dA, k, dB, n, 0.5, dC, n) == 0
    got = download(dC).reshape(m, n)
    want = 1.5 * (A.reshape(m, k).astype(np.float64) @ B.reshape(k, n)) + 0.5 * C.reshape(m, n)
    scale = 1.5 * (np.abs(A.reshape(m, k)).astype(np.float64) @ np.abs(B.reshape(k, n))) \
        + 0.5 * np.abs(C.reshape(m, n))
    assert np.all(np.abs(got - want) <= orc.fp32_tol_mult(k) * scale)


def test_odd_leading_dimensions_and_offset_pointers():
    """Odd leading dimensions and pointers one float past a 16-byte boundary:
    the 16-byte loads and stores need only dword alignment."""
    import torch
    m, n, k = 131, 67, 77
    c = Case(m, n, k, lda=m + 2, ldb=k + 1, ldc=m + 2, seed=17)
    for config in range(N_F32_CONFIGS):
        dA = torch.zeros(c.A.size + 1, dtype=torch.float32, device="cuda")
        dB = torch.zeros(c.B.size + 1, dtype=torch.float32, device="cuda")
        dC = torch.zeros(c.C0.size + 1, dtype=torch.float32, device="cuda")
        dA[1:] = upload(c.A)
        dB[1:] = upload(c.B)
        dC[1:] = upload(c.C0)
        base = 4  # one float
        log(f"unaligned config {config}")
        sa.run_config(0, config, m, n, k, c.alpha, dA.data_ptr() + base, c.lda,
                      dB.data_ptr() + base, c.ldb, c.beta, dC.data_ptr() + base, c.ldc)
        out = download(dC)[1:]
        assert_ok(check(c, out, orc.fp32_tol_mult(k)), f"unaligned config {config}")


def test_stream_argument():
    import torch
    s = torch.cuda.Stream()
    c = Case(513, 257, 129, seed=23)
    out = run_form(c, sa.FORM_F32, stream=s)
    assert_ok(check(c, out, orc.fp32_tol_mult(c.k)), "explicit stream")


# ---------------------------------------------------------------- full sizes
def _sampled_check(c: Case, out: np.ndarray, n_samples=4096, seed=0, tol_mult=None):
    rng = np.random.default_rng(seed)
    ii = rng.integers(0, c.m, n_samples).astype(np.int32)
    jj = rng.integers(0, c.n, n_samples).astype(np.int32)
    # always include the four corners and the last row/column edges
    ii = np.concatenate([ii, [0, c.m - 1, 0, c.m - 1]]).astype(np.int32)
    jj = np.concatenate([jj, [0, 0, c.n - 1, c.n - 1]]).astype(np.int32)
    ref, scale = orc.dgemm_entries(c.k, c.alpha, c.A, c.lda, c.B, c.ldb, c.beta, c.C0, c.ldc, ii, jj)
    got = out[ii + jj.astype(np.int64) * c.ldc].astype(np.float64)
    tol = (orc.fp32_tol_mult(c.k) if tol_mult is None else tol_mult) * scale
    err = np.abs(got - ref)
    assert np.all(np.isfinite(got)) and np.all(err <= tol), \
        f"{(c.m, c.n, c.k)}: max err/tol {float(np.max(err / tol))}"
    return float(np.max(err / np.maximum(scale, 1e-300)))


@pytest.mark.parametrize("m,n,k", [(1000, 1000, 1000), (2048, 2048, 2048), (4097, 3071, 2049),
                                   (8192, 8192, 127), (8192, 8192, 8192), (16384, 16384, 16384)])
def test_sweep_shapes_sampled_fp64_and_full_vs_naive(m, n, k):
    """Baseline sweep shapes: sampled FP64 entries, plus every entry against the
    independent naive kernel (itself bit-exact to the oracle's restatement)."""
    import torch
    c = Case(m, n, k, seed=m + 2 * n + 3 * k)
    dA, dB, dC = c.device()
    dC2 = dC.clone()
    log(f"sweep {(m, n, k)} sgemm")
    assert sa.sgemm(m, n, k, c.alpha, dA, c.lda, dB, c.ldb, c.beta, dC, c.ldc) == 0
    log(f"sweep {(m, n, k)} sgemm_basic")
    assert sa.sgemm_basic(m, n, k, c.alpha, dA, c.lda, dB, c.ldb, c.beta, dC2, c.ldc) == 0
    torch.cuda.synchronize()
    out = download(dC)
    _sampled_check(c, out)
    # naive kernel is within tol of truth, the fast one too: differ by <= 2 tol
    # scale = |alpha| |A||B| + |beta| |C0|, by the naive kernel (not the one under test)
    absA = dA.abs()
    absB = dB.abs()
    scale = torch.zeros_like(dC)
    assert sa.sgemm_basic(m, n, k, abs(c.alpha), absA, c.lda, absB, c.ldb, 0.0, scale, c.ldc) == 0
    scale += abs(c.beta) * dC2.new_tensor(c.C0).abs()
    diff = (dC - dC2).abs()
    bound = 2 * orc.fp32_tol_mult(k) * 1.001 * scale
    assert bool(torch.all(diff <= bound)), float((diff / bound).max())


# ------------------------------------------------------------------- split-K
SPLIT_CONFIGS = (6, 7, 8, 10)


@pytest.mark.parametrize("config", SPLIT_CONFIGS)
def test_splitk_deterministic_and_in_tolerance(config):
    """Split-K sums its partials in split order, so repeated launches give
    identical bits; results meet the FP32 tolerance against FP64.  A second
    problem of the same shape in between reuses the same partial slots, so a
    reduction that read stale partials would fail its tolerance check."""
    for (m, n, k) in [(1024, 1024, 1024), (300, 200, 5000), (129, 257, 777), (64, 64, 8192)]:
        c = Case(m, n, k, seed=m + k, alpha=1.25, beta=0.5)
        first = run_form(c, sa.FORM_F32, config)
        assert_ok(check(c, first, orc.fp32_tol_mult(k)), f"split-K {(m, n, k)} config {config}")
        other = Case(m, n, k, seed=m + k + 1000, alpha=1.0, beta=0.0)
        out = run_form(other, sa.FORM_F32, config)
        assert_ok(check(other, out, orc.fp32_tol_mult(k)), f"split-K 2nd {(m, n, k)} config {config}")
        again = run_form(c, sa.FORM_F32, config)
        assert np.array_equal(first, again), (m, n, k)


def test_splitk_graph_replays_equal_eager():
    """The per-tile counters reset themselves, so a captured split-K launch
    replays correctly many times (no memset in the graph)."""
    import torch
    m, n, k = 512, 384, 4096
    c = Case(m, n, k, seed=5, alpha=1.0, beta=0.0)
    dA, dB, dC = c.device()
    sa.run_config(0, 6, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m)  # allocates the workspace
    torch.cuda.synchronize()
    eager = download(dC).copy()
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        sa.run_config(0, 6, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, s)  # this stream's workspace
        s.synchronize()
        with torch.cuda.graph(g, stream=s):
            sa.run_config(0, 6, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, torch.cuda.current_stream())
    for rep in range(4):
        dC.fill_(123.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(download(dC), eager), rep


def test_splitk_first_call_on_busy_nonblocking_stream():
    """A stream's split-K workspace is allocated by its first split call and
    its arrival counters are zeroed on that stream, ahead of the launch (a
    null-stream memset is not ordered before work on torch's non-blocking
    streams).  Recycled device memory (a freed grouped plan's non-zero
    descriptors) and a device kept busy by another stream must not leak into
    the counts: a stale count would leave tiles of the result unwritten.
    torch hands out pooled streams, so a trial's stream is new to the library
    only while the pool has unused ones; the serving tests, whose servers
    create their own streams, cover the same path."""
    import torch
    m, n, k = 1350, 143, 1915  # 64x64 split-K tiles (config 7), six splits
    busy_c = Case(4096, 4096, 4096, seed=3, alpha=1.0, beta=0.0)
    bA, bB, bC = busy_c.device()
    for trial in range(3):
        junk = [sa.problem(300, 200, 100, 1.0, bA, 300, bB, 100, 0.0, bC, 300)] * 64
        plan = sa.GroupedPlan(256, sa.FORM_F32)
        plan.set(junk)
        plan.run()
        torch.cuda.synchronize()
        plan.destroy()
        c = Case(m, n, k, seed=40 + trial, alpha=1.25, beta=0.5)
        dA, dB, dC = c.device()
        busy, fresh = torch.cuda.Stream(), torch.cuda.Stream()
        torch.cuda.synchronize()
        sa.sgemm_ex(4096, 4096, 4096, 1.0, bA, 4096, bB, 4096, 0.0, bC, 4096, busy)
        log(f"split-K first call on a fresh stream, trial {trial}")
        sa.run_config(0, 7, m, n, k, c.alpha, dA, m, dB, k, c.beta, dC, m, fresh)
        torch.cuda.synchronize()
        assert_ok(check(c, download(dC), orc.fp32_tol_mult(k)), f"fresh stream trial {trial}")


def test_splitk_capture_without_warmup_runs_unsplit():
    """A split-K launch captured on a stream that has no workspace yet does
    not allocate inside the capture: it runs the unsplit tile (config 0's
    kernel), so each replay equals config 0's eager result bit for bit."""
    import torch
    m, n, k = 512, 384, 4096
    c = Case(m, n, k, seed=11, alpha=1.0, beta=0.0)
    dA, dB, dC = c.device()
    sa.run_config(0, 0, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m)
    torch.cuda.synchronize()
    unsplit = download(dC).copy()
    fresh = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=fresh):
        sa.run_config(0, 6, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, torch.cuda.current_stream())
    for rep in range(3):
        dC.fill_(-1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(download(dC), unsplit), rep


# ------------------------------------------------------------ tail split
TAIL_CONFIG = 9  # config 5's tile; the tiles of a partial last round split k


@pytest.mark.parametrize("m,n,k,ld_pad", [
    (3584, 3584, 1024, 0),   # 784 tiles of 128x128: one partial round of 1024 slots
    (3584, 3584, 1024, 1),   # register-staged tile (ld not a multiple of 4): 768 + 16
    (3000, 4100, 777, 5),    # ragged edges and k chunks, padded leading dimensions
    (8192, 8064, 1000, 0),   # 4032 tiles: 192 left over
])
def test_tail_split_deterministic_and_in_tolerance(m, n, k, ld_pad):
    """Whole tiles for the full rounds, split-K for the last partial round:
    in FP32 tolerance against FP64 (every entry below 2^26 of them, sampled
    above), bit-identical on repeat with another problem of the same shape
    in between (the split partial slots are reused), and every entry within
    twice the tolerance of config 5 (same tile, no split)."""
    import torch
    c = Case(m, n, k, lda=m + ld_pad, ldb=k + ld_pad, ldc=m + ld_pad, seed=m + n + k,
             alpha=1.25, beta=0.5)
    first = run_form(c, sa.FORM_F32, TAIL_CONFIG)
    if m * n <= (1 << 24):
        assert_ok(check(c, first, orc.fp32_tol_mult(k)), f"tail split {(m, n, k)}")
    else:
        _sampled_check(c, first, n_samples=4096, seed=k)
    other = Case(m, n, k, lda=m + ld_pad, ldb=k + ld_pad, ldc=m + ld_pad, seed=m + n + k + 7,
                 alpha=1.0, beta=0.0)
    out = run_form(other, sa.FORM_F32, TAIL_CONFIG)
    _sampled_check(other, out, n_samples=2048, seed=k + 1)
    again = run_form(c, sa.FORM_F32, TAIL_CONFIG)
    assert np.array_equal(first, again), (m, n, k)
    plain = run_form(c, sa.FORM_F32, 5)
    got, ref = compact(first, m, n, c.ldc), compact(plain, m, n, c.ldc)
    # |got - plain| <= |got - truth| + |truth - plain| <= 2 tol * scale, with
    # scale = |alpha| |A| |B| + |beta| |C0| from the naive kernel on the GPU
    dA, dB = upload(np.abs(c.A)), upload(np.abs(c.B))
    scale = upload(np.abs(c.C0) * abs(c.beta))
    assert sa.sgemm_basic(m, n, k, abs(c.alpha), dA, c.lda, dB, c.ldb, 1.0, scale, c.ldc) == 0
    torch.cuda.synchronize()
    bound = 2 * orc.fp32_tol_mult(k) * 1.001 * compact(download(scale), m, n, c.ldc)
    assert np.all(np.abs(got.astype(np.float64) - ref) <= bound)
    assert padding_untouched(first, c.C0, m, n, c.ldc)


def test_tail_split_graph_replays_and_capture_fallback():
    """Captured after a warm-up call on the stream, the tail split replays
    bit-identically to its eager result.  Captured on a fresh stream (no
    workspace yet), it runs config 5's unsplit kernel, so every replay
    equals config 5 bit for bit.  4096 x 4224: 1056 tiles of 128x128, one
    full round of the direct-loaded tile's 1024 slots and 32 left over."""
    import torch
    m, n = 4096, 4224
    k = 2048
    c = Case(m, n, k, seed=77, alpha=1.0, beta=0.0)
    dA, dB, dC = c.device()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        sa.run_config(0, TAIL_CONFIG, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, s)
        s.synchronize()
        eager = download(dC).copy()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            sa.run_config(0, TAIL_CONFIG, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m,
                          torch.cuda.current_stream())
    for rep in range(3):
        dC.fill_(5.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(download(dC), eager), rep
    sa.run_config(0, 5, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m)
    torch.cuda.synchronize()
    unsplit = download(dC).copy()
    assert not np.array_equal(unsplit, eager)  # the tail tiles were split
    fresh = torch.cuda.Stream()
    g2 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g2, stream=fresh):
        sa.run_config(0, TAIL_CONFIG, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m,
                      torch.cuda.current_stream())
    for rep in range(2):
        dC.fill_(-2.0)
        torch.cuda.synchronize()
        g2.replay()
        torch.cuda.synchronize()
        assert np.array_equal(download(dC), unsplit), rep


# ------------------------------------------------------------ stream-K
STREAMK_CONFIG = 11  # config 5's tile; stream-K over a partial round's tiles


@pytest.mark.parametrize("m,n,k,ld_pad", [
    (1000, 1100, 2001, 3),   # 72 tiles of 128x128, no full round; ragged last k-tile
    (2816, 2816, 1536, 0),   # 484 tiles
    (3584, 3584, 1024, 1),   # 784 tiles: 16 past one round
])
def test_streamk_deterministic_and_in_tolerance(m, n, k, ld_pad):
    """In FP32 tolerance against FP64 (every entry below 2^21 of them,
    sampled above), bit-identical on repeat with another problem of the same
    shape in between (the pieces' partial slots and counters are reused)."""
    c = Case(m, n, k, lda=m + ld_pad, ldb=k + ld_pad, ldc=m + ld_pad, seed=3 * m + n + k,
             alpha=1.25, beta=0.5)
    first = run_form(c, sa.FORM_F32, STREAMK_CONFIG)
    if m * n <= (1 << 21):
        assert_ok(check(c, first, orc.fp32_tol_mult(k)), f"stream-K {(m, n, k)}")
    else:
        _sampled_check(c, first, n_samples=4096, seed=k)
    assert padding_untouched(first, c.C0, m, n, c.ldc)
    other = Case(m, n, k, lda=m + ld_pad, ldb=k + ld_pad, ldc=m + ld_pad, seed=m + n + k + 13,
                 alpha=1.0, beta=0.0)
    out = run_form(other, sa.FORM_F32, STREAMK_CONFIG)
    _sampled_check(other, out, n_samples=2048, seed=k + 5)
    again = run_form(c, sa.FORM_F32, STREAMK_CONFIG)
    assert np.array_equal(first, again), (m, n, k)


def test_streamk_integer_kat_equals_whole_tiles():
    """Integer inputs: every partial sum is an exact integer, so stream-K
    equals config 5 (whole tiles) bit for bit."""
    m, n, k = 2816, 2432, 1536
    c = Case(m, n, k, kind="int", seed=29, alpha=1.5, beta=0.5)
    assert np.array_equal(run_form(c, sa.FORM_F32, 5), run_form(c, sa.FORM_F32, STREAMK_CONFIG))


def test_streamk_graph_replays_and_capture_fallback():
    """Captured after a warm-up call on the stream, stream-K replays
    bit-identically to its eager result; captured on a fresh stream (no
    workspace yet) it runs config 5's kernel on whole tiles, bit for bit."""
    import torch
    m, n, k = 2816, 2560, 1024
    c = Case(m, n, k, seed=81, alpha=1.0, beta=0.0)
    dA, dB, dC = c.device()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        sa.run_config(0, STREAMK_CONFIG, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, s)
        s.synchronize()
        eager = download(dC).copy()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            sa.run_config(0, STREAMK_CONFIG, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m,
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
        sa.run_config(0, STREAMK_CONFIG, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m,
                      torch.cuda.current_stream())
    for rep in range(2):
        dC.fill_(-2.0)
        torch.cuda.synchronize()
        g2.replay()
        torch.cuda.synchronize()
        assert np.array_equal(download(dC), whole), rep


def test_streamk_concurrent_streams_match_sequential():
    """Each stream has its own stream-K workspace: two stream-K launches of
    different problems running at the same time on two streams give the bits
    each gives alone."""
    import torch
    shapes = [(2816, 2560, 1536), (2432, 2600, 1024)]
    cases = [Case(m, n, k, seed=95 + i, alpha=1.0, beta=0.0) for i, (m, n, k) in enumerate(shapes)]
    devs = [c.device() for c in cases]
    alone = []
    for c, (dA, dB, dC) in zip(cases, devs):
        sa.run_config(0, STREAMK_CONFIG, c.m, c.n, c.k, 1.0, dA, c.m, dB, c.k, 0.0, dC, c.m)
        torch.cuda.synchronize()
        alone.append(download(dC).copy())
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for rep in range(3):
        for (_, _, dC) in devs:
            dC.fill_(3.0)
        torch.cuda.synchronize()
        for c, (dA, dB, dC), st in zip(cases, devs, streams):
            sa.run_config(0, STREAMK_CONFIG, c.m, c.n, c.k, 1.0, dA, c.m, dB, c.k, 0.0, dC,
                          c.m, st)
        torch.cuda.synchronize()
        for i, (_, _, dC) in enumerate(devs):
            assert np.array_equal(download(dC), alone[i]), (rep, i)


@pytest.mark.parametrize("form", [sa.FORM_F32, sa.FORM_BF16X3])
def test_result_larger_than_4gib(form):
    """C of 65539 x 16389 floats (4.3 GB): element offsets pass 2^32 bytes, so
    the epilogue's 64-bit addressing and the per-k-tile buffer rebasing are
    exercised; beta = 0 over a NaN-filled C also checks that C is not read.
    Sampled entries (corners included) against FP64."""
    import torch
    m, n, k = 65539, 16389, 67
    lda, ldb, ldc = m + 1, k, m
    A = orc.fill("uniform", orc.colmajor_size(m, k, lda), 901)
    B = orc.fill("uniform", orc.colmajor_size(k, n, ldb), 902)
    dA, dB = upload(A), upload(B)
    dC = torch.full((ldc * n,), float("nan"), device="cuda", dtype=torch.float32)
    log(f"4GiB C form {form}")
    fn = sa.sgemm_ex if form == sa.FORM_F32 else sa.sgemm_bf16x3
    fn(m, n, k, 1.0, dA, lda, dB, ldb, 0.0, dC, ldc)
    rng = np.random.default_rng(5)
    ii = np.concatenate([rng.integers(0, m, 4000), [0, m - 1, 0, m - 1]]).astype(np.int64)
    jj = np.concatenate([rng.integers(0, n, 4000), [0, 0, n - 1, n - 1]]).astype(np.int64)
    idx = torch.from_numpy(ii + jj * ldc).to("cuda")
    got = dC[idx].double().cpu().numpy()
    ref, scale = orc.dgemm_entries(k, 1.0, A, lda, B, ldb, 0.0, None, ldc, ii, jj)
    tol = (orc.fp32_tol_mult(k) if form == sa.FORM_F32 else orc.bf16x3_tol_mult(k)) * scale
    assert np.all(np.isfinite(got))
    assert np.all(np.abs(got - ref) <= tol), float(np.max(np.abs(got - ref) / tol))
    del dC
    torch.cuda.empty_cache()


@pytest.mark.parametrize("form", [sa.FORM_F32, sa.FORM_BF16X3])
def test_huge_leading_dimension_fallback(form):
    """lda or ldb above 2^21 (past the tiles' 32-bit buffer offsets): the
    default selection runs the wide fallback (naive FP32, 64-bit addressing,
    beta = 0 does not read C); results within the FP32 tolerance."""
    big = (1 << 21) + 3
    fn = sa.sgemm_ex if form == sa.FORM_F32 else sa.sgemm_bf16x3
    for (m, n, k, lda, ldb) in [(40, 3, 5, big, 5), (7, 3, 33, 7, big)]:
        c = Case(m, n, k, lda=lda, ldb=ldb, seed=m + k, alpha=1.5, beta=0.0, c_fill=np.nan)
        dA, dB, dC = c.device()
        log(f"wide ld form {form} {(m, n, k, lda, ldb)}")
        fn(m, n, k, c.alpha, dA, c.lda, dB, c.ldb, c.beta, dC, c.ldc)
        out = download(dC)
        assert_ok(check(c, out, orc.fp32_tol_mult(k)), f"wide ld {(m, n, k, lda, ldb)}")


def test_concurrent_host_threads_on_separate_streams():
    """Entry points are reentrant: four host threads, each on its own stream,
    run split-K and tiled problems at once (each stream gets its own split-K
    workspace); every result equals the one-thread result bit for bit."""
    import threading
    import torch
    shapes = [(1024, 1024, 1024), (256, 256, 8192), (8192, 64, 4096), (3000, 500, 700)]
    cases = [Case(m, n, k, seed=11 + i, alpha=1.0, beta=0.5) for i, (m, n, k) in enumerate(shapes)]
    dev = [c.device() for c in cases]
    ref = []
    for c, (dA, dB, dC) in zip(cases, dev):
        out = dC.clone()
        sa.sgemm_ex(c.m, c.n, c.k, c.alpha, dA, c.lda, dB, c.ldb, c.beta, out, c.ldc)
        ref.append(download(out))
    outs = [[None] * 5 for _ in cases]
    errors = []

    def worker(i):
        try:
            c, (dA, dB, dC) = cases[i], dev[i]
            s = torch.cuda.Stream()
            for r in range(5):
                out = dC.clone()
                torch.cuda.synchronize()
                sa.sgemm_ex(c.m, c.n, c.k, c.alpha, dA, c.lda, dB, c.ldb, c.beta, out, c.ldc, s)
                s.synchronize()
                outs[i][r] = out.cpu().numpy()
        except Exception as e:  # reported below, in the main thread
            errors.append(repr(e))

    log("concurrent threads")
    threads = [threading.Thread(target=worker, args=(i,)) for i in range(len(cases))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(120)
    assert not errors, errors
    for i in range(len(cases)):
        for r in range(5):
            assert np.array_equal(outs[i][r], ref[i]), (shapes[i], r)
