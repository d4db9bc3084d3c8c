"""Non-finite and near-overflow inputs, both forms, against the FP64 oracle.

IEEE semantics as the reference's FP32 kernels have them (src/kernels/
128x128x8.cuh: plain FFMA, so Inf * b = +-Inf, Inf - Inf = NaN, NaN
propagates, and beta * NaN reaches C when beta != 0):
  * where the FP64 oracle's entry is NaN, the result is NaN;
  * where it is +-Inf, the result is the same infinity;
  * every finite entry meets the form's stated tolerance.
The FP64 oracle sees the same IEEE operations, and whether an entry is NaN
or +-Inf does not depend on the summation order here: no finite partial sum
overflows (the huge entries meet B rows scaled by 1/4).

BF16x3: hi is rounded from x clamped to the largest finite bf16, so +-Inf in
an operand lands in lo and reaches C through lo * b_hi with the sign of the
FP32 product, and finite values above the largest bf16 (3.3961e38 ..
FLT_MAX, which would round to Inf) split exactly.  Inherent limit of the
two-term split, not tested: a result within about 2^-17 of FLT_MAX can
overflow (hi + lo of FLT_MAX itself rounds to 2^128).
"""
import numpy as np
import pytest

import oracle as orc
import sgemm_amd as sa
from gpu_util import Case, compact, download, log, upload

pytestmark = pytest.mark.gpu


def _case(m, n, k, seed):
    c = Case(m, n, k, lda=m + 3, ldb=k + 1, ldc=m + 2, seed=seed)
    lda, ldb, ldc = c.lda, c.ldb, c.ldc
    A, B, C0 = c.A, c.B, c.C0
    A[3 + 5 * lda] = np.inf
    A[10 + 7 * lda] = -np.inf
    A[20 + 9 * lda] = np.nan
    B[11 + 4 * ldb] = np.inf          # column 4: +-Inf in every row with A[i, 11] != 0
    B[13 + 6 * ldb] = -np.inf
    C0[30 + 2 * ldc] = np.nan          # beta != 0: NaN reaches C[30, 2]
    # row 40: finite values above the largest bf16 (3.3895e38), below FLT_MAX
    # by far more than 2^-17; B rows 0 and 1 scaled so no FP32 sum overflows
    A[40 + 0 * lda] = np.float32(3.397e38)
    A[40 + 1 * lda] = np.float32(-3.401e38)
    A[41 + 2 * lda] = np.float32(3.3962e38)
    for j in range(n):
        B[0 + j * ldb] *= np.float32(0.25)
        B[1 + j * ldb] *= np.float32(0.25)
        B[2 + j * ldb] *= np.float32(0.25)
    return c


def _check(c, out, tol_mult, what):
    ref, scale = c.reference()
    got = compact(out, c.m, c.n, c.ldc).astype(np.float64)
    assert np.isnan(ref).sum() > 0 and np.isinf(ref).sum() > 0, "case must exercise both"
    nan_ref, inf_ref = np.isnan(ref), np.isinf(ref)
    assert np.array_equal(np.isnan(got), nan_ref), f"{what}: NaN pattern differs"
    assert np.array_equal(got[inf_ref], ref[inf_ref]), f"{what}: infinities differ"
    fin = np.isfinite(ref)
    d = np.abs(got[fin] - ref[fin])
    bad = ~(d <= tol_mult * scale[fin])
    assert not bad.any(), f"{what}: {int(bad.sum())} finite entries outside tolerance"
    # the huge rows stay finite and are computed, not clamped (columns 4 and 6
    # carry B's infinities into every row)
    cols = [j for j in range(c.n) if j not in (4, 6)]
    grid = compact(out, c.m, c.n, c.ldc).reshape(c.n, c.m)
    for row in (40, 41):
        v = grid[cols, row]
        assert np.isfinite(v).all() and np.abs(v).max() > 1e37, (what, row)


FORMS = [(sa.FORM_F32, orc.fp32_tol_mult), (sa.FORM_BF16X3, orc.bf16x3_tol_mult)]


@pytest.mark.parametrize("form,tolf", FORMS)
def test_nonfinite_every_config(form, tolf):
    m, n, k = 300, 257, 1100
    c = _case(m, n, k, seed=4242)
    dA, dB = upload(c.A), upload(c.B)
    for cfg in range(len(sa.configs(form))):
        log(f"nonfinite form={form} config={cfg} {m}x{n}x{k}")
        dC = upload(c.C0)
        sa.run_config(form, cfg, m, n, k, c.alpha, dA, c.lda, dB, c.ldb, c.beta, dC, c.ldc)
        _check(c, download(dC), tolf(k), f"form {form} config {cfg}")


@pytest.mark.parametrize("form,tolf", FORMS)
@pytest.mark.parametrize("shape", [(2048, 1536, 2048), (700, 900, 4096)])
def test_nonfinite_default_entry(form, tolf, shape):
    """The default entry points (the split-K / stream-K picks included)."""
    m, n, k = shape
    c = _case(m, n, k, seed=m + n)
    dA, dB, dC = c.device()
    log(f"nonfinite default form={form} {m}x{n}x{k}")
    fn = sa.sgemm_ex if form == sa.FORM_F32 else sa.sgemm_bf16x3
    fn(m, n, k, c.alpha, dA, c.lda, dB, c.ldb, c.beta, dC, c.ldc)
    _check(c, download(dC), tolf(k), f"form {form} default {shape}")
