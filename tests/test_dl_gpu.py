"""The FP32 tiles with direct-to-LDS operand loads: config 5's 128x128x16
tile (configs 5, 9 and 10) and config 1/7's 64x64x32 tile run them when A, B,
lda and ldb give 16-byte aligned columns; otherwise the register-staged tile
runs.  Both tiles do the same arithmetic in the same
order, so the same matrices give the same bits whichever one runs: each case
here runs the aligned layout (direct loads) and the same values at ld + 1 or
at a 4-byte offset (register staging), compares the bits, and checks the
aligned result against the FP64 oracle.  The k values cover one partial
k-tile (k below the tile's 16 or 32, register path only), exactly one and two
full k-tiles, a partial last k-tile after the direct-loaded ones, and long
k."""
import numpy as np
import pytest

import oracle as orc
import sgemm_amd as sa
from gpu_util import Case, assert_ok, check, compact, download, log, upload

pytestmark = pytest.mark.gpu

SHAPES = [(128, 128, 16), (256, 128, 32), (130, 257, 33), (300, 257, 129), (129, 130, 1000),
          (64, 64, 8192), (1000, 1000, 1000), (127, 385, 7), (513, 260, 47), (64, 192, 64),
          (200, 72, 95)]


def _aligned(v):
    return (v + 3) // 4 * 4


def _relayout(x, rows, cols, ld, new_ld, offset=0):
    """Column-major rows x cols values of an ld-strided buffer, re-strided
    to new_ld after `offset` leading floats (filled with a marker)."""
    out = np.full(offset + orc.colmajor_size(rows, cols, new_ld), -7.0, dtype=np.float32)
    for j in range(cols):
        out[offset + j * new_ld: offset + j * new_ld + rows] = x[j * ld: j * ld + rows]
    return out


def _run(config, c, A, lda, B, ldb, a_off=0, b_off=0):
    dA, dB, dC = upload(A), upload(B), upload(c.C0)
    log(f"dl config={config} m={c.m} n={c.n} k={c.k} lda={lda} ldb={ldb} "
        f"a_off={a_off} b_off={b_off}")
    a_ptr = dA.data_ptr() + 4 * a_off
    b_ptr = dB.data_ptr() + 4 * b_off
    st = sa.run_config(sa.FORM_F32, config, c.m, c.n, c.k, c.alpha, a_ptr, lda, b_ptr, ldb,
                       c.beta, dC, c.ldc, None)
    assert st == 0, st
    out = download(dC)
    del dA, dB
    return out


@pytest.mark.parametrize("config", [1, 5, 7, 9, 10])
def test_direct_loads_match_register_staging(config):
    for (m, n, k) in SHAPES:
        lda, ldb = _aligned(m), _aligned(k)
        c = Case(m, n, k, lda=lda, ldb=ldb, ldc=m + 1, seed=3 * m + n + k, alpha=1.5, beta=0.5)
        direct = _run(config, c, c.A, lda, c.B, ldb)
        assert_ok(check(c, direct, orc.fp32_tol_mult(k)), f"dl {(m, n, k)} config {config}")
        # the same values with ld + 1: register-staged tile
        A1 = _relayout(c.A, m, k, lda, lda + 1)
        B1 = _relayout(c.B, k, n, ldb, ldb + 1)
        staged = _run(config, c, A1, lda + 1, B1, ldb + 1)
        assert np.array_equal(compact(direct, m, n, c.ldc), compact(staged, m, n, c.ldc)), \
            f"ld + 1 {(m, n, k)} config {config}"
        # aligned leading dimensions, operands 4 bytes past a 16-byte boundary
        A2 = _relayout(c.A, m, k, lda, lda, offset=1)
        B2 = _relayout(c.B, k, n, ldb, ldb, offset=1)
        shifted = _run(config, c, A2, lda, B2, ldb, a_off=1, b_off=1)
        assert np.array_equal(compact(direct, m, n, c.ldc), compact(shifted, m, n, c.ldc)), \
            f"offset pointers {(m, n, k)} config {config}"


@pytest.mark.parametrize("config", [1, 5, 7, 9, 10])
def test_direct_loads_integer_kat_bit_exact(config):
    """Integer inputs in [-4, 4] with aligned columns: exact in any order."""
    for (m, n, k) in [(300, 256, 129), (132, 128, 1000), (64, 64, 8192), (128, 128, 16)]:
        c = Case(m, n, k, lda=_aligned(m), ldb=_aligned(k), ldc=m + 5, kind="int",
                 seed=m + n + k)
        out = _run(config, c, c.A, c.lda, c.B, c.ldb)
        ref, _ = c.reference()
        assert np.array_equal(compact(out, m, n, c.ldc).astype(np.float64), ref), (m, n, k)


def test_direct_loads_tail_split_and_splitk_at_scale():
    """Configs 9 (tail split), 10 and 7 (split-K) with the direct-loaded tiles
    on grids that split: bit-identical to config 5 on integer data, and config
    5 equal to the FP64 oracle on sampled entries (integer data is exact)."""
    for (m, n, k, config) in [(3584, 3584, 1024, 9), (2048, 1024, 4096, 10),
                              (8192, 8064, 1008, 9), (1024, 512, 8192, 7)]:
        c = Case(m, n, k, kind="int", seed=m + k, alpha=1.5, beta=0.5)
        split = _run(config, c, c.A, c.lda, c.B, c.ldb)
        whole = _run(5, c, c.A, c.lda, c.B, c.ldb)
        assert np.array_equal(split, whole), (m, n, k, config)
        rng = np.random.default_rng(k)
        idx = rng.integers(0, m * n, size=2048)
        rows, cols = idx % m, idx // m
        ref = orc.dgemm_entries(k, c.alpha, c.A, c.lda, c.B, c.ldb, c.beta, c.C0, c.ldc,
                                rows, cols)[0]
        got = whole[cols * c.ldc + rows].astype(np.float64)
        assert np.array_equal(got, ref), (m, n, k)
