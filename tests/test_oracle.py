"""Pins the CPU oracle (oracle/oracle.c) to values that do not come from it.

Each oracle function is checked against at least one of: the published
SplitMix64 stream, the golden fixtures of tests/golden/ (exact rational
arithmetic), numpy's float64 matmul, closed forms, or hand-worked BF16
rounding cases.  The checks are chosen so that a dropped term, a wrong sign or
index, or a transposed operand fails.
"""
import json
import math
import os
import struct

import numpy as np
import pytest

import oracle as orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def cases():
    return _load("gemm_cases.json")["cases"]


# ------------------------------------------------------------ SplitMix64
def test_splitmix64_published_stream():
    # First outputs for seed 1234567 as published with the reference
    # implementation of SplitMix64 (Vigna, splitmix64.c).
    assert orc.splitmix64_stream(1234567, 5) == [
        6457827717110365317, 3203168211198807973, 9817491932198370423,
        4593380528125082431, 16408922859458223821]
    # seed 0: 0xE220A8397B1DCDAF is the widely quoted first output.
    assert orc.splitmix64_stream(0, 1) == [0xE220A8397B1DCDAF]


def test_splitmix64_matches_fixture_and_counter_form():
    streams = _load("splitmix64.json")["streams"]
    L = orc.lib()
    for seed, vals in streams.items():
        seed = int(seed)
        assert orc.splitmix64_stream(seed, len(vals)) == [int(v) for v in vals]
        assert [L.orc_splitmix64_at(seed, i) for i in range(len(vals))] == [int(v) for v in vals]


def test_fill_uniform_definition():
    streams = _load("splitmix64.json")["streams"]
    z = [int(v) for v in streams["1"]]
    x = orc.fill("uniform", len(z), 1)
    want = np.array([2.0 * ((v >> 40) / 2.0 ** 24) - 1.0 for v in z], dtype=np.float32)
    assert np.array_equal(x, want)
    big = orc.fill("uniform", 1 << 16, 7)
    assert big.min() >= -1.0 and big.max() < 1.0
    assert np.all(big * 2.0 ** 23 == np.round(big * 2.0 ** 23))  # 24-bit grid
    assert abs(float(big.mean())) < 0.02


def test_fill_int_and_wide():
    x = orc.fill("int", 1 << 14, 3, -4, 4)
    assert set(np.unique(x).tolist()) == set(float(v) for v in range(-4, 5))
    w = orc.fill("wide", 1 << 14, 5)
    e = np.log2(np.abs(w.astype(np.float64)))
    assert e.min() >= -8.0 - 1e-6 and e.max() < 8.0 + 1e-6
    assert 0.4 < float((w < 0).mean()) < 0.6


# ------------------------------------------------------------- FP64 truth
def test_dgemm_golden_exact(cases):
    for c in cases:
        ref, scale = orc.dgemm(c["m"], c["n"], c["k"], c["alpha"], np.array(c["A"]), c["lda"],
                               np.array(c["B"]), c["ldb"], c["beta"], np.array(c["C0"]), c["ldc"])
        exp = np.array(c["expected_f64"])
        # FP64 summation of <= 31 exact products of 24-bit values: error tiny.
        assert np.allclose(ref, exp, rtol=1e-14, atol=1e-14), c["name"]
        if c["kind"] == "int":
            assert np.array_equal(ref, exp), c["name"]  # integers: exact


def _rand_cm(rng, rows, cols, ld):
    buf = rng.uniform(-1, 1, size=ld * (cols - 1) + rows).astype(np.float32)
    return buf


def _dense(buf, rows, cols, ld):
    return np.stack([buf[j * ld: j * ld + rows] for j in range(cols)], axis=1).astype(np.float64)


@pytest.mark.parametrize("m,n,k,lda,ldb,ldc", [(7, 5, 3, 7, 3, 7), (33, 17, 65, 40, 70, 35),
                                                 (64, 1, 128, 64, 128, 64), (1, 50, 9, 3, 9, 2)])
def test_dgemm_vs_numpy(m, n, k, lda, ldb, ldc):
    rng = np.random.default_rng(m * 1000 + n)
    A, B, C = _rand_cm(rng, m, k, lda), _rand_cm(rng, k, n, ldb), _rand_cm(rng, m, n, ldc)
    alpha, beta = -1.25, 0.75
    ref, scale = orc.dgemm(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
    Ad, Bd, Cd = _dense(A, m, k, lda), _dense(B, k, n, ldb), _dense(C, m, n, ldc)
    want = alpha * (Ad @ Bd) + beta * Cd
    assert np.allclose(ref.reshape(n, m).T, want, rtol=1e-12, atol=1e-12)
    want_scale = abs(alpha) * (np.abs(Ad) @ np.abs(Bd)) + abs(beta) * np.abs(Cd)
    assert np.allclose(scale.reshape(n, m).T, want_scale, rtol=1e-12, atol=1e-12)
    # a transposed operand would not match (non-square shapes make it fail)
    if m != k:
        assert not np.allclose(ref.reshape(n, m).T[: min(m, k)], 0)


def test_dgemm_beta_zero_does_not_read_c():
    m, n, k = 9, 4, 6
    rng = np.random.default_rng(1)
    A, B = _rand_cm(rng, m, k, m), _rand_cm(rng, k, n, k)
    C = np.full(m * n, np.nan, dtype=np.float32)
    ref, scale = orc.dgemm(m, n, k, 2.0, A, m, B, k, 0.0, C, m)
    assert np.all(np.isfinite(ref)) and np.all(np.isfinite(scale))


def test_dgemm_closed_forms():
    m, n, k = 6, 5, 11
    ones_a = np.ones(m * k, np.float32)
    ones_b = np.ones(k * n, np.float32)
    C = np.arange(m * n, dtype=np.float32)
    ref, _ = orc.dgemm(m, n, k, 1.5, ones_a, m, ones_b, k, 0.5, C, m)
    assert np.array_equal(ref, 1.5 * k + 0.5 * C.astype(np.float64))
    # identity * B == B ; A * permutation permutes A's columns
    eye = np.eye(k, dtype=np.float32).reshape(-1, order="F").copy()
    B = np.arange(k * n, dtype=np.float32)
    ref, _ = orc.dgemm(k, n, k, 1.0, eye, k, B, k, 0.0, None, k)
    assert np.array_equal(ref, B.astype(np.float64))
    perm = np.array([2, 0, 1])
    P = np.zeros((3, 3), np.float32)
    P[perm, np.arange(3)] = 1.0  # column j of A*P is column perm[j] of A
    A = np.arange(4 * 3, dtype=np.float32)
    ref, _ = orc.dgemm(4, 3, 3, 1.0, A, 4, P.reshape(-1, order="F").copy(), 3, 0.0, None, 4)
    Ad = A.reshape(3, 4).T
    assert np.array_equal(ref.reshape(3, 4).T, Ad[:, perm])


def test_dgemm_k_zero_gives_beta_c():
    C = np.arange(12, dtype=np.float32)
    ref, _ = orc.dgemm(3, 4, 0, 1.0, np.zeros(1, np.float32), 3, np.zeros(1, np.float32), 1,
                       0.5, C, 3)
    assert np.array_equal(ref, 0.5 * C)


def test_dgemm_entries_match_full(cases):
    c = cases[2]
    ref, scale = orc.dgemm(c["m"], c["n"], c["k"], c["alpha"], np.array(c["A"]), c["lda"],
                           np.array(c["B"]), c["ldb"], c["beta"], np.array(c["C0"]), c["ldc"])
    ii = np.array([0, 5, c["m"] - 1, 3], dtype=np.int32)
    jj = np.array([0, 2, c["n"] - 1, 7], dtype=np.int32)
    out, sc = orc.dgemm_entries(c["k"], c["alpha"], np.array(c["A"]), c["lda"], np.array(c["B"]),
                                c["ldb"], c["beta"], np.array(c["C0"]), c["ldc"], ii, jj)
    assert np.array_equal(out, ref[ii + jj * c["m"]])
    assert np.array_equal(sc, scale[ii + jj * c["m"]])


# ----------------------------------------------------- FP32 restatement
def test_sgemm_basic_bit_exact_vs_golden(cases):
    """kernel_basic's order (fma chain, *alpha, fma with beta) rounded exactly
    step by step in tests/golden/make_golden.py must equal the C restatement."""
    for c in cases:
        C = np.array(c["C0"], dtype=np.float32)
        orc.sgemm_basic_cm(c["m"], c["n"], c["k"], c["alpha"], np.array(c["A"]), c["lda"],
                           np.array(c["B"]), c["ldb"], c["beta"], C, c["ldc"])
        got = np.array([C[i + j * c["ldc"]] for j in range(c["n"]) for i in range(c["m"])],
                       dtype=np.float32)
        want = np.array(c["expected_basic_f32"], dtype=np.float32)
        assert np.array_equal(got, want), c["name"]
        # padding rows of C untouched
        for j in range(c["n"] - 1):
            for i in range(c["m"], c["ldc"]):
                assert C[i + j * c["ldc"]] == 777.0


def test_sgemm_basic_rowmajor_is_colmajor_swap():
    rng = np.random.default_rng(5)
    m, n, k = 13, 7, 9
    A = rng.uniform(-1, 1, m * k).astype(np.float32)   # row-major m x k, lda = k
    B = rng.uniform(-1, 1, k * n).astype(np.float32)   # row-major k x n, ldb = n
    C = rng.uniform(-1, 1, m * n).astype(np.float32)   # row-major m x n, ldc = n
    C_rm = C.copy()
    orc.sgemm_basic_rm(m, n, k, 1.5, A, k, B, n, 0.5, C_rm, n)
    want = 1.5 * (A.reshape(m, k).astype(np.float64) @ B.reshape(k, n)) + 0.5 * C.reshape(m, n)
    assert np.allclose(C_rm.reshape(m, n), want, rtol=1e-5, atol=1e-5)
    # The same memory as a column-major (n x m) product with operands swapped
    C_cm = C.copy()
    orc.lib().orc_sgemm_basic_rm(m, n, k, 1.5, A, k, B, n, 0.5, C_cm, n)
    assert np.array_equal(C_cm, C_rm)


def test_sgemm_basic_always_reads_c():
    # kernel_basic computes acc + C*beta even for beta == 0: NaN propagates.
    C = np.full(4, np.nan, dtype=np.float32)
    orc.sgemm_basic_cm(2, 2, 1, 1.0, np.ones(2, np.float32), 2, np.ones(2, np.float32), 1, 0.0,
                       C, 2)
    assert np.all(np.isnan(C))


# ------------------------------------------------------------------ BF16
def test_bf16_rne_hand_cases():
    for c in _load("bf16_rne.json")["cases"]:
        x = struct.unpack("<f", struct.pack("<I", c["f32_bits"]))[0]
        got = orc.bf16_rne(x)
        if math.isnan(x):
            assert (got & 0x7F80) == 0x7F80 and (got & 0x7F) != 0, c["what"]
        else:
            assert got == c["bf16_bits"], (c["what"], hex(got))


def test_bf16x3_split_properties():
    x = orc.fill("wide", 4096, 9)
    hi, lo = orc.bf16x3_split(x)
    h = np.array([orc.bf16_to_f32(int(v)) for v in hi], dtype=np.float64)
    l = np.array([orc.bf16_to_f32(int(v)) for v in lo], dtype=np.float64)
    xd = x.astype(np.float64)
    assert np.all(np.abs(xd - h) <= 2.0 ** -8 * np.abs(xd))         # hi within half an ulp
    assert np.all(np.abs(xd - h - l) <= 2.0 ** -16 * np.abs(xd))    # residual after lo
    # hi exact for values with <= 8 significant bits
    y = np.array([1.0, -3.0, 0.5, 2.0 ** -20, 96.0], np.float32)
    hy, ly = orc.bf16x3_split(y)
    assert np.all(ly == 0)


def test_bf16x3_split_non_finite_and_overflow():
    """hi is rounded from x clamped to +-0x7f7f (the largest finite bf16,
    (2 - 2^-7) * 2^127), so it never overflows.  Expected bits worked by hand:
      +Inf -> hi 0x7f7f, lo = bf16(Inf - hi) = +Inf 0x7f80 (and the negatives);
      FLT_MAX = (2 - 2^-23) 2^127 -> hi 0x7f7f, residual 2^120 - 2^104, which
        rounds to 2^120: lo 0x7b80 (exponent 247, mantissa 0);
      0x7f7f8000 (the tie above 0x7f7f, plain RNE goes to Inf) -> hi 0x7f7f,
        residual exactly 2^119: lo 0x7b00;
      NaN -> lo is a NaN."""
    def bits(x):
        return np.array([struct.unpack("<f", struct.pack("<I", b))[0] for b in x], np.float32)
    x = bits([0x7F800000, 0xFF800000, 0x7F7FFFFF, 0xFF7FFFFF, 0x7F7F8000, 0x7F7F0000])
    hi, lo = orc.bf16x3_split(x)
    assert [int(v) for v in hi] == [0x7F7F, 0xFF7F, 0x7F7F, 0xFF7F, 0x7F7F, 0x7F7F]
    assert [int(v) for v in lo] == [0x7F80, 0xFF80, 0x7B80, 0xFB80, 0x7B00, 0x0000]
    # hi + lo reproduces every finite clamped value to 2^-16 relative
    for xv, h, l in zip(x[2:], hi[2:], lo[2:]):
        s = orc.bf16_to_f32(int(h)) + orc.bf16_to_f32(int(l))
        assert abs(s - float(xv)) <= 2.0 ** -16 * abs(float(xv))
    _, ln = orc.bf16x3_split(np.array([np.nan], np.float32))
    assert (int(ln[0]) & 0x7F80) == 0x7F80 and (int(ln[0]) & 0x7F) != 0


# ------------------------------------------------------ metric / tolerance
def test_gamma_and_tolerance():
    u = 2.0 ** -24
    assert orc.gamma(1) == pytest.approx(u / (1 - u), rel=1e-15)
    assert orc.gamma(8192) == pytest.approx(8192 * u / (1 - 8192 * u), rel=1e-15)
    assert orc.fp32_tol_mult(1024) == pytest.approx(2 * orc.gamma(1026))
    assert orc.bf16x3_tol_mult(64) > orc.fp32_tol_mult(64)


def test_check_counts_planted_errors():
    m, n, k = 8, 6, 16
    rng = np.random.default_rng(3)
    A, B = _rand_cm(rng, m, k, m), _rand_cm(rng, k, n, k)
    ref, scale = orc.dgemm(m, n, k, 1.0, A, m, B, k, 0.0, None, m)
    C = ref.astype(np.float32)
    tol = orc.fp32_tol_mult(k)
    st = orc.check(m, n, C, m, ref, scale, tol)
    assert st["n_fail"] == 0 and st["max_ratio"] <= 1.0
    bad = C.copy()
    bad[5] += 10 * tol * scale[5] + 1e-3
    bad[7] = np.nan
    bad[9] = np.inf
    st = orc.check(m, n, bad, m, ref, scale, tol)
    assert st["n_fail"] == 3 and st["n_nan"] == 1 and st["n_inf"] == 1
    # a dropped term (one product missing) is far outside the tolerance
    A2 = A.copy()
    A2[0] = 0.0
    ref2, _ = orc.dgemm(m, n, k, 1.0, A2, m, B, k, 0.0, None, m)
    st = orc.check(m, n, ref2.astype(np.float32), m, ref, scale, tol)
    assert st["n_fail"] > 0


def test_fnv1a64_known_values():
    # FNV-1a 64-bit of "a" and of "foobar" (published test vectors).
    assert orc.fnv1a64(np.frombuffer(b"a", dtype=np.uint8)) == 0xAF63DC4C8601EC8C
    assert orc.fnv1a64(np.frombuffer(b"foobar", dtype=np.uint8)) == 0x85944171F73967E8
