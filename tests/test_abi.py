"""The C ABI surface, checked without a GPU: the library loads, exports every
entry point include/sgemm.h declares, and rejects invalid arguments on the
host before any HIP call (the reference validates nothing,
src/sgemm.cuh:19-90)."""
import ctypes
import os
import subprocess

import pytest

import sgemm_amd as sa


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(sa.LIB_PATH):
        sa.build()
    return sa.lib()


def test_header_symbols_exported(lib):
    names = sa.header_symbols()
    assert "sgemm" in names and "sgemm_basic" in names and "sgemm_ex" in names
    out = subprocess.run(["nm", "-D", "--defined-only", sa.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in names if n not in exported]
    assert not missing, f"declared in include/sgemm.h but not exported: {missing}"
    for n in names:
        assert hasattr(lib, n)


def test_invalid_arguments_rejected_on_host(lib):
    # Every case fails validation before the library touches the GPU.
    p = 4096  # never dereferenced
    cases = [
        (-1, 4, 4, 4, 4, 4), (4, -1, 4, 4, 4, 4), (4, 4, -1, 4, 4, 4),
        (4, 4, 4, 3, 4, 4), (4, 4, 4, 4, 3, 4), (4, 4, 4, 4, 4, 3), (0, 0, 0, 0, 1, 1),
    ]
    for (m, n, k, lda, ldb, ldc) in cases:
        assert sa.sgemm(m, n, k, 1.0, p, lda, p, ldb, 0.0, p, ldc) == 1
        assert sa.sgemm_ex(m, n, k, 1.0, p, lda, p, ldb, 0.0, p, ldc, check=False) == 1
        assert sa.sgemm_bf16x3(m, n, k, 1.0, p, lda, p, ldb, 0.0, p, ldc, check=False) == 1
    assert sa.sgemm(4, 4, 4, 1.0, None, 4, p, 4, 0.0, p, 4) == 1
    assert sa.sgemm(4, 4, 4, 1.0, p, 4, p, 4, 0.0, None, 4) == 1
    # operands that are not 4-byte aligned (no float pointer is): rejected
    for A, B, C in ((p + 2, p, p), (p, p + 1, p), (p, p, p + 3)):
        assert sa.sgemm_ex(4, 4, 4, 1.0, A, 4, B, 4, 0.0, C, 4, check=False) == 1
        assert sa.sgemm_bf16x3(4, 4, 4, 1.0, A, 4, B, 4, 0.0, C, 4, check=False) == 1
        assert sa.sgemm_ex_noalloc(0, 4, 4, 4, 1.0, A, 4, B, 4, 0.0, C, 4, check=False) == 1
    # null alpha / beta pointers
    L = sa.lib()
    assert L.sgemm_ex(4, 4, 4, None, p, 4, p, 4, None, p, 4, None) == 1
    # the void entry point records its status once, then clears it
    L.sgemm(-1, 1, 1, None, None, 1, None, 1, None, None, 1)
    assert L.sgemm_get_last_error() == 1
    assert L.sgemm_get_last_error() == 0


def test_empty_problems_are_noops_without_gpu(lib):
    # m == 0 or n == 0 returns before any launch (no device needed).
    assert sa.sgemm(0, 8, 8, 1.0, None, 1, None, 8, 0.0, None, 1) == 0
    assert sa.sgemm(8, 0, 8, 1.0, None, 8, None, 8, 0.0, None, 8) == 0


def test_config_queries(lib):
    for form in (sa.FORM_F32, sa.FORM_BF16X3):
        for c in sa.configs(form):
            assert c["bm"] > 0 and c["bn"] > 0 and c["bk"] > 0 and c["name"]
    assert len(sa.configs(sa.FORM_F32)) >= 1
    assert lib.sgemm_config_info(7, 0, None, None, None, None, 0) == 2
    assert sa.status_string(5).startswith("request exceeds")
    assert "gfx950" in sa.version()


def test_huge_leading_dimension_explicit_config_not_supported(lib):
    """An explicit tile config cannot address leading dimensions above 2^21
    (its buffer offsets are 32-bit): rejected on the host, nothing launched.
    The default selection runs such problems on the wide fallback kernel
    (tests/test_sgemm_gpu.py::test_huge_leading_dimension_fallback)."""
    p = 4096
    big = (1 << 21) + 1
    assert sa.run_config(0, 0, 4, 4, 4, 1.0, p, big, p, 4, 0.0, p, 4, check=False) == 2
    assert sa.run_config(1, 2, 4, 4, 4, 1.0, p, 4, p, big, 0.0, p, 4, check=False) == 2


def test_unknown_config_rejected(lib):
    p = 4096
    assert sa.run_config(0, 999, 4, 4, 4, 1.0, p, 4, p, 4, 0.0, p, 4, check=False) == 2
    assert sa.run_config(5, 0, 4, 4, 4, 1.0, p, 4, p, 4, 0.0, p, 4, check=False) == 2
