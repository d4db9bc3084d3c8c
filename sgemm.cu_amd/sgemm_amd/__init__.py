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


class Problem(ctypes.Structure):
    """Mirror of sgemm_problem (include/sgemm.h)."""

    _fields_ = [
        ("m", ctypes.c_int), ("n", ctypes.c_int), ("k", ctypes.c_int),
        ("alpha", ctypes.c_float), ("beta", ctypes.c_float),
        ("A", ctypes.c_void_p), ("lda", ctypes.c_int),
        ("B", ctypes.c_void_p), ("ldb", ctypes.c_int),
        ("C", ctypes.c_void_p), ("ldc", ctypes.c_int),
    ]


EXECUTOR_GPU = 0
EXECUTOR_MOCK = 1
OPERANDS_DEVICE = 0
OPERANDS_HOST = 1


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_size_t), ("max_batch", ctypes.c_int),
        ("num_slots", ctypes.c_int), ("form", ctypes.c_int), ("executor", ctypes.c_int),
        ("mock_tflops", ctypes.c_double), ("mock_launch_us", ctypes.c_double),
        ("chunk_gflop", ctypes.c_double), ("mock_fail_launch", ctypes.c_int),
    ]


class Request(ctypes.Structure):
    _fields_ = [
        ("m", ctypes.c_int), ("n", ctypes.c_int), ("k", ctypes.c_int),
        ("alpha", ctypes.c_float), ("beta", ctypes.c_float),
        ("A", ctypes.c_void_p), ("lda", ctypes.c_int),
        ("B", ctypes.c_void_p), ("ldb", ctypes.c_int),
        ("C", ctypes.c_void_p), ("ldc", ctypes.c_int),
        ("operands", ctypes.c_int),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_longlong), ("completed", ctypes.c_longlong),
        ("rejected", ctypes.c_longlong), ("failed", ctypes.c_longlong),
        ("batches", ctypes.c_longlong), ("measured", ctypes.c_longlong),
        ("mean_batch", ctypes.c_double), ("p50_ms", ctypes.c_double),
        ("p99_ms", ctypes.c_double), ("mean_ms", ctypes.c_double), ("max_ms", ctypes.c_double),
        ("mean_queue_ms", ctypes.c_double), ("window_s", ctypes.c_double),
        ("tflops", ctypes.c_double), ("peak_bytes", ctypes.c_size_t),
        ("budget_bytes", ctypes.c_size_t), ("panel_launches", ctypes.c_longlong),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


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
    # One HIP runtime per process.  torch's ROCm build ships its own
    # libamdhip64 and asks for it by the unversioned name, so if this library
    # loaded /opt/rocm's libamdhip64.so.7 first, importing torch afterwards
    # would bring a second HIP/HSA runtime into the process (no device seen by
    # one of them, crashes in the other).  Loading torch first makes this
    # library bind to torch's runtime through its SONAME instead.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
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
    L.sgemm_ex_noalloc.argtypes = [i] + mm + [vp]
    L.sgemm_ex_noalloc.restype = i
    L.sgemm_run_config.argtypes = [i, i] + mm + [vp]
    L.sgemm_run_config.restype = i
    L.sgemm_config_count.argtypes = [i]
    L.sgemm_config_count.restype = i
    L.sgemm_config_info.argtypes = [i, i, ctypes.POINTER(i), ctypes.POINTER(i),
                                    ctypes.POINTER(i), ctypes.c_char_p, i]
    L.sgemm_config_info.restype = i
    L.sgemm_version.restype = ctypes.c_char_p
    L.sgemm_hip_error_string.restype = ctypes.c_char_p
    P = ctypes.POINTER(Problem)
    plan_p = ctypes.c_void_p
    for name, args in {
        "sgemm_grouped_plan_create": [ctypes.POINTER(plan_p), i, i],
        "sgemm_grouped_plan_set": [plan_p, i, P, vp],
        "sgemm_grouped_plan_set_configs": [plan_p, i, P, ctypes.POINTER(i), vp],
        "sgemm_grouped_plan_run": [plan_p, vp],
        "sgemm_grouped_plan_capture": [plan_p, vp],
        "sgemm_grouped_plan_replay": [plan_p, vp],
        "sgemm_grouped_plan_info": [plan_p, ctypes.POINTER(i), ctypes.POINTER(i),
                                    ctypes.POINTER(i), ctypes.POINTER(ctypes.c_double)],
        "sgemm_grouped_plan_destroy": [plan_p],
        "sgemm_grouped": [i, P, i, vp],
        "sgemm_autotune_bucket": [i, i, i, i, i, i, vp, ctypes.POINTER(i),
                                  ctypes.POINTER(ctypes.c_float)],
        "sgemm_autotune_lookup": [i, i, i, i, i, ctypes.POINTER(i), ctypes.POINTER(i)],
        "sgemm_launch_config": [i, i, i, i, vp, i, vp, i, ctypes.POINTER(i)],
        "sgemm_autotune_set": [i, i, i, i, i, i],
        "sgemm_autotune_clear": [],
        "sgemm_autotune_size": [],
        "sgemm_autotune_save": [ctypes.c_char_p],
        "sgemm_autotune_load": [ctypes.c_char_p],
        "sgemm_server_create": [ctypes.POINTER(plan_p), ctypes.POINTER(ServerConfig)],
        "sgemm_server_submit": [plan_p, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)],
        "sgemm_server_poll": [plan_p, ctypes.c_uint64],
        "sgemm_server_wait": [plan_p, ctypes.c_uint64, ctypes.c_double],
        "sgemm_server_drain": [plan_p, ctypes.c_double],
        "sgemm_server_latency": [plan_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(ctypes.c_double)],
        "sgemm_server_get_stats": [plan_p, ctypes.POINTER(ServerStats)],
        "sgemm_server_reset_stats": [plan_p],
        "sgemm_server_destroy": [plan_p],
        "sgemm_workspace_release": [vp],
    }.items():
        getattr(L, name).argtypes = args
        getattr(L, name).restype = i
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
        if status == 3:
            where += f" [{lib().sgemm_hip_error_string().decode()}]"
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


def sgemm_ex_noalloc(form, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream=None,
                     check=True) -> int:
    """sgemm_ex / sgemm_bf16x3 with scratch from a static device array (no allocation)."""
    st = lib().sgemm_ex_noalloc(form, m, n, k, _f(alpha), _ptr(A), lda, _ptr(B), ldb, _f(beta),
                                _ptr(C), ldc, _stream(stream))
    if check:
        _check(st, "sgemm_ex_noalloc")
    return st


def run_config(form, config_id, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream=None,
               check=True) -> int:
    st = lib().sgemm_run_config(form, config_id, m, n, k, _f(alpha), _ptr(A), lda, _ptr(B), ldb,
                                _f(beta), _ptr(C), ldc, _stream(stream))
    if check:
        _check(st, "sgemm_run_config")
    return st


def workspace_release(stream=None) -> None:
    """Free the library's split-K workspace and packed BF16x3 planes of `stream`
    (sgemm_workspace_release)."""
    _check(lib().sgemm_workspace_release(_stream(stream)), "sgemm_workspace_release")


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


# ------------------------------------------------------------ grouped GEMM
def problem(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc) -> Problem:
    return Problem(m, n, k, alpha, beta, _ptr(A), lda, _ptr(B), ldb, _ptr(C), ldc)


def _problem_array(problems):
    arr = (Problem * max(1, len(problems)))()
    for idx, p in enumerate(problems):
        arr[idx] = p
    return arr


class GroupedPlan:
    """A grouped-GEMM plan: descriptors for up to `capacity` problems, one
    persistent launch per run, optional hipGraph capture/replay."""

    def __init__(self, capacity: int, form: int = FORM_F32):
        self.handle = ctypes.c_void_p()
        _check(lib().sgemm_grouped_plan_create(ctypes.byref(self.handle), capacity, form),
               "sgemm_grouped_plan_create")

    def set(self, problems, stream=None, configs=None) -> None:
        arr = _problem_array(problems)
        if configs is None:
            _check(lib().sgemm_grouped_plan_set(self.handle, len(problems), arr, _stream(stream)),
                   "sgemm_grouped_plan_set")
        else:
            cfg = (ctypes.c_int * max(1, len(configs)))(*configs)
            _check(lib().sgemm_grouped_plan_set_configs(self.handle, len(problems), arr, cfg,
                                                        _stream(stream)),
                   "sgemm_grouped_plan_set_configs")

    def run(self, stream=None) -> None:
        _check(lib().sgemm_grouped_plan_run(self.handle, _stream(stream)), "sgemm_grouped_plan_run")

    def capture(self, stream) -> None:
        _check(lib().sgemm_grouped_plan_capture(self.handle, _stream(stream)),
               "sgemm_grouped_plan_capture")

    def replay(self, stream=None) -> None:
        _check(lib().sgemm_grouped_plan_replay(self.handle, _stream(stream)),
               "sgemm_grouped_plan_replay")

    def info(self) -> dict:
        c, t, g = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        f = ctypes.c_double()
        _check(lib().sgemm_grouped_plan_info(self.handle, ctypes.byref(c), ctypes.byref(t),
                                             ctypes.byref(g), ctypes.byref(f)),
               "sgemm_grouped_plan_info")
        return {"count": c.value, "total_tiles": t.value, "grid": g.value, "flops": f.value}

    def destroy(self) -> None:
        if self.handle:
            _check(lib().sgemm_grouped_plan_destroy(self.handle), "sgemm_grouped_plan_destroy")
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            if self.handle:
                lib().sgemm_grouped_plan_destroy(self.handle)
        except Exception:
            pass


def grouped(problems, form=FORM_F32, stream=None) -> None:
    arr = _problem_array(problems)
    _check(lib().sgemm_grouped(len(problems), arr, form, _stream(stream)), "sgemm_grouped")


# --------------------------------------------------------------- autotuner
def autotune_bucket(form, m, n, k, iters=3, grouped=False, stream=None):
    cfg, tf = ctypes.c_int(), ctypes.c_float()
    _check(lib().sgemm_autotune_bucket(form, int(grouped), m, n, k, iters, _stream(stream),
                                       ctypes.byref(cfg), ctypes.byref(tf)),
           "sgemm_autotune_bucket")
    return cfg.value, tf.value


def autotune_lookup(form, m, n, k, grouped=False):
    cfg, from_table = ctypes.c_int(), ctypes.c_int()
    _check(lib().sgemm_autotune_lookup(form, int(grouped), m, n, k, ctypes.byref(cfg),
                                       ctypes.byref(from_table)), "sgemm_autotune_lookup")
    return cfg.value, bool(from_table.value)


def launch_config(form, m, n, k, A, lda, B, ldb) -> int:
    """Config the default entry point launches for these operands (-1: none)."""
    cfg = ctypes.c_int()
    _check(lib().sgemm_launch_config(form, m, n, k, _ptr(A), lda, _ptr(B), ldb,
                                     ctypes.byref(cfg)), "sgemm_launch_config")
    return cfg.value


def autotune_set(form, m, n, k, config_id, grouped=False):
    _check(lib().sgemm_autotune_set(form, int(grouped), m, n, k, config_id), "sgemm_autotune_set")


def autotune_clear():
    _check(lib().sgemm_autotune_clear(), "sgemm_autotune_clear")


def autotune_size() -> int:
    return int(lib().sgemm_autotune_size())


def autotune_save(path: str):
    _check(lib().sgemm_autotune_save(path.encode()), "sgemm_autotune_save")


def autotune_load(path: str):
    _check(lib().sgemm_autotune_load(path.encode()), "sgemm_autotune_load")


# ------------------------------------------------------------ serving loop
class Server:
    """Continuous-batching GEMM server (sgemm_server_* in include/sgemm.h)."""

    def __init__(self, budget_bytes, max_batch=64, num_slots=2, form=FORM_BF16X3,
                 executor=EXECUTOR_GPU, mock_tflops=0.0, mock_launch_us=0.0, chunk_gflop=0.0,
                 mock_fail_launch=0):
        cfg = ServerConfig(budget_bytes, max_batch, num_slots, form, executor, mock_tflops,
                           mock_launch_us, chunk_gflop, mock_fail_launch)
        self.handle = ctypes.c_void_p()
        _check(lib().sgemm_server_create(ctypes.byref(self.handle), ctypes.byref(cfg)),
               "sgemm_server_create")

    def submit(self, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, operands=OPERANDS_DEVICE,
               check=True):
        r = Request(m, n, k, alpha, beta, _hptr(A), lda, _hptr(B), ldb, _hptr(C), ldc, operands)
        rid = ctypes.c_uint64()
        st = lib().sgemm_server_submit(self.handle, ctypes.byref(r), ctypes.byref(rid))
        if check:
            _check(st, "sgemm_server_submit")
        return rid.value if st == 0 else -st

    def poll(self, rid) -> int:
        return lib().sgemm_server_poll(self.handle, rid)

    def wait(self, rid, timeout=60.0) -> int:
        return lib().sgemm_server_wait(self.handle, rid, timeout)

    def drain(self, timeout=120.0) -> None:
        _check(lib().sgemm_server_drain(self.handle, timeout), "sgemm_server_drain")

    def latency(self, rid):
        t, q = ctypes.c_double(), ctypes.c_double()
        _check(lib().sgemm_server_latency(self.handle, rid, ctypes.byref(t), ctypes.byref(q)),
               "sgemm_server_latency")
        return t.value, q.value

    def stats(self) -> dict:
        st = ServerStats()
        _check(lib().sgemm_server_get_stats(self.handle, ctypes.byref(st)), "sgemm_server_stats")
        return st.as_dict()

    def reset_stats(self) -> None:
        _check(lib().sgemm_server_reset_stats(self.handle), "sgemm_server_reset_stats")

    def destroy(self) -> None:
        if self.handle:
            _check(lib().sgemm_server_destroy(self.handle), "sgemm_server_destroy")
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            if self.handle:
                lib().sgemm_server_destroy(self.handle)
        except Exception:
            pass


def _hptr(x):
    """Device tensor / int pointer, or a numpy array (host operands)."""
    if x is None:
        return None
    if hasattr(x, "ctypes") and hasattr(x, "dtype"):
        return x.ctypes.data
    return _ptr(x)
