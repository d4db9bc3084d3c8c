"""Helpers shared by the GPU parity tests: seeded inputs from the oracle's
SplitMix64 generator, upload through torch (device memory only), a call
through the C ABI, and the FP64 oracle check with the stated tolerance."""
from __future__ import annotations

import sys

import numpy as np

import oracle as orc
import sgemm_amd as sa


def torch_dev():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("GPU test needs a HIP device")
    return torch


def log(msg: str) -> None:
    """Name the case before each GPU step (flushed, so a fault still shows it)."""
    print(msg, file=sys.stderr, flush=True)


def upload(x: np.ndarray):
    torch = torch_dev()
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to("cuda")


def download(t) -> np.ndarray:
    torch = torch_dev()
    torch.cuda.synchronize()
    return t.cpu().numpy()


class Case:
    """Seeded column-major operands for one (m, n, k, lda, ldb, ldc) problem."""

    def __init__(self, m, n, k, lda=None, ldb=None, ldc=None, kind="uniform", seed=1,
                 alpha=1.5, beta=0.5, c_fill=None):
        self.m, self.n, self.k = m, n, k
        self.lda = lda if lda is not None else max(1, m)
        self.ldb = ldb if ldb is not None else max(1, k)
        self.ldc = ldc if ldc is not None else max(1, m)
        self.alpha, self.beta = alpha, beta
        self.kind = kind
        self.A = orc.fill(kind, orc.colmajor_size(m, k, self.lda), seed)
        self.B = orc.fill(kind, orc.colmajor_size(k, n, self.ldb), seed + 1)
        if c_fill is None:
            self.C0 = orc.fill(kind, orc.colmajor_size(m, n, self.ldc), seed + 2)
        else:
            self.C0 = np.full(orc.colmajor_size(m, n, self.ldc), c_fill, dtype=np.float32)

    def reference(self):
        return orc.dgemm(self.m, self.n, self.k, self.alpha, self.A, self.lda, self.B, self.ldb,
                         self.beta, self.C0, self.ldc)

    def device(self):
        return upload(self.A), upload(self.B), upload(self.C0)


def run_form(case: Case, form: int, config: int | None = None, stream=None):
    """Run one problem; returns the downloaded C buffer (ldc-strided)."""
    dA, dB, dC = case.device()
    c = case
    log(f"run form={form} config={config} m={c.m} n={c.n} k={c.k} lda={c.lda} ldb={c.ldb} "
        f"ldc={c.ldc} alpha={c.alpha} beta={c.beta} kind={c.kind}")
    if config is None:
        fn = sa.sgemm_ex if form == sa.FORM_F32 else sa.sgemm_bf16x3
        fn(c.m, c.n, c.k, c.alpha, dA, c.lda, dB, c.ldb, c.beta, dC, c.ldc, stream)
    else:
        sa.run_config(form, config, c.m, c.n, c.k, c.alpha, dA, c.lda, dB, c.ldb, c.beta, dC,
                      c.ldc, stream)
    return download(dC)


def compact(buf: np.ndarray, m: int, n: int, ld: int) -> np.ndarray:
    """Column-major (m x n) values of an ld-strided buffer, as a 1-D array."""
    if m == 0 or n == 0:
        return np.zeros(0, dtype=buf.dtype)
    idx = (np.arange(m)[:, None] + ld * np.arange(n)[None, :]).reshape(-1, order="F")
    return buf[idx]


def padding_untouched(out: np.ndarray, before: np.ndarray, m: int, n: int, ld: int) -> bool:
    """Rows m..ld-1 of every column but the last must be unchanged."""
    if ld == m or n <= 1:
        return True
    mask = np.ones(out.size, dtype=bool)
    mask[(np.arange(m)[:, None] + ld * np.arange(n)[None, :]).reshape(-1)] = False
    return bool(np.array_equal(out[mask], before[mask], equal_nan=True))


def check(case: Case, out: np.ndarray, tol_mult: float, ref=None) -> dict:
    if ref is None:
        ref = case.reference()
    r, s = ref
    got = compact(out, case.m, case.n, case.ldc)
    st = orc.check(case.m, case.n, got, case.m, r, s, tol_mult)
    st["padding_ok"] = padding_untouched(out, case.C0, case.m, case.n, case.ldc)
    return st


def assert_ok(st: dict, what: str) -> None:
    assert st["n_fail"] == 0 and st["padding_ok"], f"{what}: {st}"
