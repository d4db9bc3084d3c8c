"""Operands whose leading dimension is a multiple of 4 (the 16-byte load
path) while the row or k count is not: the last column of A (or B) ends in
the middle of a 16-byte group, so one load per lane reaches past the end of
the operand.  The values below the end must still arrive intact: a load that
straddles the end of a buffer range may be cut as a whole, which would zero
valid rows of the last column."""
import numpy as np
import pytest

import oracle as orc
import sgemm_amd as sa
from gpu_util import Case, assert_ok, check, run_form

pytestmark = pytest.mark.gpu

SHAPES = [(130, 64, 66), (129, 70, 127), (131, 33, 5), (258, 129, 1025), (67, 300, 35)]


@pytest.mark.parametrize("form", [sa.FORM_F32, sa.FORM_BF16X3])
def test_vector_path_with_ragged_last_column(form):
    tol = orc.fp32_tol_mult if form == sa.FORM_F32 else orc.bf16x3_tol_mult
    for (m, n, k) in SHAPES:
        lda, ldb, ldc = (m + 3) // 4 * 4, (k + 3) // 4 * 4, (m + 3) // 4 * 4
        if lda == m:
            lda += 4
        if ldb == k:
            ldb += 4
        for cfg in range(len(sa.configs(form))):
            c = Case(m, n, k, lda=lda, ldb=ldb, ldc=ldc, seed=m + 7 * n + cfg)
            out = run_form(c, form, cfg)
            assert_ok(check(c, out, tol(k)), f"form {form} config {cfg} {(m, n, k)}")
