"""Python binding of the CPU oracle (liboracle.so, built from oracle/oracle.c).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only as the checker (or, for the CPU
baseline, as the timed host restatement of the reference's kernel_basic,
src/kernels/basic.cuh:15-25).  The product library never loads it.

All matrices are numpy arrays in column-major storage: a logical (rows, cols)
matrix with leading dimension ld is a 1-D float32 buffer of at least
ld*(cols-1)+rows elements, element (i, j) at i + j*ld.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Stats(ctypes.Structure):
    """Mirror of orc_stats in oracle.c."""

    _fields_ = [
        ("max_err", ctypes.c_double),
        ("mean_err", ctypes.c_double),
        ("max_rel", ctypes.c_double),
        ("mean_rel", ctypes.c_double),
        ("max_ratio", ctypes.c_double),
        ("n_fail", ctypes.c_longlong),
        ("n_nan", ctypes.c_longlong),
        ("n_inf", ctypes.c_longlong),
        ("n_checked", ctypes.c_longlong),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    args = ["make", "-C", _HERE]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = ctypes.CDLL(_LIB_PATH)
    c_int, c_double, c_size, c_u64 = ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_uint64
    L.orc_splitmix64.argtypes = [ctypes.POINTER(c_u64)]
    L.orc_splitmix64.restype = c_u64
    L.orc_splitmix64_at.argtypes = [c_u64, c_u64]
    L.orc_splitmix64_at.restype = c_u64
    L.orc_fill_uniform.argtypes = [_f32p, c_size, c_u64]
    L.orc_fill_wide.argtypes = [_f32p, c_size, c_u64]
    L.orc_fill_int.argtypes = [_f32p, c_size, c_u64, c_int, c_int]
    L.orc_fnv1a64.argtypes = [ctypes.c_void_p, c_size]
    L.orc_fnv1a64.restype = c_u64
    L.orc_dgemm_cm.argtypes = [c_int, c_int, c_int, c_double, _f32p, c_int, _f32p, c_int,
                               c_double, ctypes.c_void_p, c_int, _f64p, ctypes.c_void_p, c_int]
    L.orc_dgemm_entries_cm.argtypes = [c_int, c_double, _f32p, c_int, _f32p, c_int, c_double,
                                       ctypes.c_void_p, c_int, c_int, _i32p, _i32p, _f64p,
                                       ctypes.c_void_p]
    L.orc_sgemm_basic_rm.argtypes = [c_int, c_int, c_int, ctypes.c_float, _f32p, c_int, _f32p,
                                     c_int, ctypes.c_float, _f32p, c_int]
    L.orc_sgemm_basic_cm.argtypes = L.orc_sgemm_basic_rm.argtypes
    L.orc_bf16_rne.argtypes = [ctypes.c_float]
    L.orc_bf16_rne.restype = ctypes.c_uint16
    L.orc_bf16_to_f32.argtypes = [ctypes.c_uint16]
    L.orc_bf16_to_f32.restype = ctypes.c_float
    L.orc_bf16x3_split.argtypes = [_f32p, c_size, _u16p, _u16p]
    L.orc_gamma.argtypes = [c_double]
    L.orc_gamma.restype = c_double
    L.orc_check.argtypes = [c_int, c_int, _f32p, c_int, _f64p, _f64p, c_int, c_double,
                            ctypes.POINTER(Stats)]
    L.orc_max_threads.restype = c_int
    _lib = L
    return L


# ---------------------------------------------------------------- generation
def splitmix64_stream(seed: int, count: int) -> list[int]:
    state = ctypes.c_uint64(seed)
    return [lib().orc_splitmix64(ctypes.byref(state)) for _ in range(count)]


def fill(kind: str, n: int, seed: int, lo: int = -4, hi: int = 4) -> np.ndarray:
    """n float32 values: kind is 'uniform' ([-1,1)), 'wide' (+-2^[-8,8)) or 'int'."""
    x = np.empty(max(n, 1), dtype=np.float32)
    if kind == "uniform":
        lib().orc_fill_uniform(x, n, seed)
    elif kind == "wide":
        lib().orc_fill_wide(x, n, seed)
    elif kind == "int":
        lib().orc_fill_int(x, n, seed, lo, hi)
    else:
        raise ValueError(f"unknown fill kind {kind!r}")
    return x[:n] if n else x[:0]


def fnv1a64(a: np.ndarray) -> int:
    a = np.ascontiguousarray(a)
    return int(lib().orc_fnv1a64(a.ctypes.data, a.nbytes))


def colmajor_size(rows: int, cols: int, ld: int) -> int:
    """Elements a column-major (rows, cols, ld) operand spans."""
    if rows <= 0 or cols <= 0:
        return 0
    return ld * (cols - 1) + rows


# ---------------------------------------------------------------- FP64 truth
def dgemm(m, n, k, alpha, A, lda, B, ldb, beta, C0, ldc, with_scale=True):
    """FP64 column-major alpha*A*B + beta*C0 -> (ref, scale), both (ldo=m) 1-D."""
    ldo = max(m, 1)
    out = np.zeros(ldo * max(n, 1), dtype=np.float64)
    scale = np.zeros_like(out) if with_scale else None
    C0keep = _as_f32(C0) if (C0 is not None and beta != 0.0) else None
    lib().orc_dgemm_cm(m, n, k, float(alpha), _as_f32(A), lda, _as_f32(B), ldb, float(beta),
                       C0keep.ctypes.data if C0keep is not None else None, ldc, out,
                       scale.ctypes.data if scale is not None else None, ldo)
    return out, scale


def dgemm_entries(k, alpha, A, lda, B, ldb, beta, C0, ldc, ii, jj):
    ii = np.ascontiguousarray(ii, dtype=np.int32)
    jj = np.ascontiguousarray(jj, dtype=np.int32)
    out = np.zeros(len(ii), dtype=np.float64)
    scale = np.zeros(len(ii), dtype=np.float64)
    C0keep = _as_f32(C0) if (C0 is not None and beta != 0.0) else None
    lib().orc_dgemm_entries_cm(k, float(alpha), _as_f32(A), lda, _as_f32(B), ldb, float(beta),
                               C0keep.ctypes.data if C0keep is not None else None, ldc,
                               len(ii), ii, jj, out, scale.ctypes.data)
    return out, scale


# ---------------------------------------------------------- FP32 restatement
def sgemm_basic_rm(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    """In-place FP32 restatement of kernel_basic (row-major). C is modified."""
    lib().orc_sgemm_basic_rm(m, n, k, alpha, _as_f32(A), lda, _as_f32(B), ldb, beta, C, ldc)


def sgemm_basic_cm(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    """Column-major view of the same restatement (operands swapped)."""
    lib().orc_sgemm_basic_cm(m, n, k, alpha, _as_f32(A), lda, _as_f32(B), ldb, beta, C, ldc)


# ------------------------------------------------------------------- BF16
def bf16_rne(x: float) -> int:
    return int(lib().orc_bf16_rne(float(x)))


def bf16_to_f32(h: int) -> float:
    return float(lib().orc_bf16_to_f32(h))


def bf16x3_split(x: np.ndarray):
    x = _as_f32(x)
    hi = np.empty(x.size, dtype=np.uint16)
    lo = np.empty(x.size, dtype=np.uint16)
    lib().orc_bf16x3_split(x, x.size, hi, lo)
    return hi, lo


# ------------------------------------------------------------- tolerances
def gamma(k: float) -> float:
    return float(lib().orc_gamma(float(k)))


def fp32_tol_mult(k: int, safety: float = 2.0) -> float:
    """Stated plain-FP32 tolerance factor: tol_ij = 2*gamma_{k+2} * scale_ij."""
    return safety * gamma(k + 2)


def bf16x3_tol_mult(k: int, safety: float = 2.0) -> float:
    """BF16x3 tolerance: the FP32 bound widened by the dropped split terms,
    3*2^-16 per product (|lo*lo| and the two residual products)."""
    return safety * (gamma(k + 2) + 3.0 * 2.0 ** -16)


def check(m, n, C, ldc, ref, scale, tol_mult) -> dict:
    st = Stats()
    lib().orc_check(m, n, _as_f32(C), ldc, np.ascontiguousarray(ref, dtype=np.float64),
                    np.ascontiguousarray(scale, dtype=np.float64), max(m, 1), float(tol_mult),
                    ctypes.byref(st))
    return st.as_dict()


def max_threads() -> int:
    return int(lib().orc_max_threads())


def _as_f32(x) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.float32).reshape(-1)
    if a.size == 0:
        a = np.zeros(1, dtype=np.float32)
    return a
