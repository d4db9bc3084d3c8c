"""Python mirror of the sgemm C ABI (include/sgemm.h) over libsgemm_amd.so.

The reference's interface is one C++ call (src/sgemm.cuh:19-30):

    sgemm(m, n, k, &alpha, A, lda, B, ldb, &beta, C, ldc)

and its test harness compares it with sgemm_basic (src/sgemm.cuh:92-113,
test.cu:78-79).  This module exposes the same names with the same argument
meaning -- column-major, device operands, host scalars -- so callers and tests
read like the reference's.  Operands are device pointers (int) or CUDA/HIP torch
tensors (their data_ptr() is used); torch is only plumbing for device memory.

The library is the product path: if libsgemm_amd.so is missing this module
raises on import of the library instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os
import re

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(_PKG_DIR)                     # sgemm.cu_amd/
LIB_PATH = os.path.join(ROOT, "lib", "libsgemm_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(ROOT), "include", "sgemm.h")

FORM_F32 = 0
FORM_BF16X3 = 1

STATUS = {
    0: "SUCCESS", 1: "INVALID_VALUE", 2: "NOT_SUPPORTED", 3: "HIP_ERROR",
    4: "ALLOC_FAILED", 5: "BUDGET_EXCEEDED", 6: "NOT_READY", 7: "NOT_FOUND",
    8: "IO_ERROR", 9: "WRONG_STATE",
}


class SgemmError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        super().__init__(f"{where}: status {status} ({STATUS.get(status, '?')})")


_lib = None


def build(force: bool = False) -> str:
    import subprocess
    args = ["make", "-C", ROOT, "-j8"]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    """Load libsgemm_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libsgemm_amd.so not found at {LIB_PATH}: build it first "
                           f"(make -C {ROOT})")
    L = ctypes.CDLL(LIB_PATH)
    i, f, vp = ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.c_void_p
    mm = [i, i, i, f, vp, i, vp, i, f, vp, i]
    for name in ("sgemm", "sgemm_basic", "sgemm_rm"):
        getattr(L, name).argtypes = mm
        getattr(L, name).restype = None
    L.sgemm_get_last_error.restype = i
    L.sgemm_status_string.argtypes = [i]
    L.sgemm_status_string.restype = ctypes.c_char_p
    L.sgemm_ex.argtypes = mm + [vp]
    L.sgemm_ex.restype = i
    L.sgemm_bf16x3.argtypes = mm + [vp]
    L.sgemm_bf16x3.restype = i
    L.sgemm_run_config.argtypes = [i, i] + mm + [vp]
    L.sgemm_run_config.restype = i
    L.sgemm_config_count.argtypes = [i]
    L.sgemm_config_count.restype = i
    L.sgemm_config_info.argtypes = [i, i, ctypes.POINTER(i), ctypes.POINTER(i),
                                    ctypes.POINTER(i), ctypes.c_char_p, i]
    L.sgemm_config_info.restype = i
    L.sgemm_version.restype = ctypes.c_char_p
    _lib = L
    return L


def header_symbols(path: str = HEADER_PATH) -> list[str]:
    """Function names declared in include/sgemm.h."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(",
                       text, flags=re.M)
    return sorted(set(n for n in names if n.startswith("sgemm")))


def _ptr(x) -> int | None:
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return int(x.data_ptr())


def _stream(s) -> int | None:
    if s is None:
        return None
    if isinstance(s, int):
        return s
    return int(s.cuda_stream)


def _f(x: float):
    return ctypes.byref(ctypes.c_float(x))


def _check(status: int, where: str) -> None:
    if status != 0:
        raise SgemmError(status, where)


# ------------------------------------------------- drop-in (reference) calls
def sgemm(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc) -> int:
    """Column-major C := alpha*A*B + beta*C on the null stream.  Returns (and
    clears) the recorded status, like a checkCudaErrors-style caller would."""
    L = lib()
    L.sgemm(m, n, k, _f(alpha), _ptr(A), lda, _ptr(B), ldb, _f(beta), _ptr(C), ldc)
    return L.sgemm_get_last_error()


def sgemm_basic(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc) -> int:
    L = lib()
    L.sgemm_basic(m, n, k, _f(alpha), _ptr(A), lda, _ptr(B), ldb, _f(beta), _ptr(C), ldc)
    return L.sgemm_get_last_error()


def sgemm_rm(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc) -> int:
    L = lib()
    L.sgemm_rm(m, n, k, _f(alpha), _ptr(A), lda, _ptr(B), ldb, _f(beta), _ptr(C), ldc)
    return L.sgemm_get_last_error()


# --------------------------------------------- status-returning, on a stream
def sgemm_ex(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream=None, check=True) -> int:
    st = lib().sgemm_ex(m, n, k, _f(alpha), _ptr(A), lda, _ptr(B), ldb, _f(beta), _ptr(C), ldc,
                        _stream(stream))
    if check:
        _check(st, "sgemm_ex")
    return st


def sgemm_bf16x3(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream=None, check=True) -> int:
    st = lib().sgemm_bf16x3(m, n, k, _f(alpha), _ptr(A), lda, _ptr(B), ldb, _f(beta), _ptr(C),
                            ldc, _stream(stream))
    if check:
        _check(st, "sgemm_bf16x3")
    return st


def run_config(form, config_id, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream=None,
               check=True) -> int:
    st = lib().sgemm_run_config(form, config_id, m, n, k, _f(alpha), _ptr(A), lda, _ptr(B), ldb,
                                _f(beta), _ptr(C), ldc, _stream(stream))
    if check:
        _check(st, "sgemm_run_config")
    return st


def configs(form: int) -> list[dict]:
    L = lib()
    out = []
    for cid in range(L.sgemm_config_count(form)):
        bm, bn, bk = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        name = ctypes.create_string_buffer(64)
        _check(L.sgemm_config_info(form, cid, ctypes.byref(bm), ctypes.byref(bn),
                                   ctypes.byref(bk), name, 64), "sgemm_config_info")
        out.append({"id": cid, "bm": bm.value, "bn": bn.value, "bk": bk.value,
                    "name": name.value.decode()})
    return out


def status_string(status: int) -> str:
    return lib().sgemm_status_string(status).decode()


def version() -> str:
    return lib().sgemm_version().decode()
