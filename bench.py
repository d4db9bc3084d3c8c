#!/usr/bin/env python3
"""Benchmark of the column-major SGEMM hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 8192]

One step = one C := alpha*A*B + beta*C call at M=N=K=size (default 8192, the
shape the headline metric is quoted on) with alpha=1.5, beta=0.5 and entries
uniform in [-1, 1) (synthetic, seeded; resident in HBM before timing starts).
The flop convention is the reference's 2*M*N*K (benchmark.cu:138).

Prints ONE JSON line.  `value` is the plain FP32 form's TFLOPS; the BF16x3
form, its speedup and the error of both forms on wide-range inputs ride
along under "forms" (8192^3) and "forms_large" (16384^3): over the whole
matrix against an FP64 matmul on the GPU, and on 8192 fixed entries against
FP64 dot products computed on the host.  "error_target" says per size
whether BF16x3 reaches 2x the plain TFLOPS at a max error no larger than the
plain form's.  "vendor" times the vendor BLAS (hipBLASLt, rocBLAS) on the
same problem, "sweep" both forms over the square sizes 1024..16384 and three
odd shapes, "grouped" the 512-problem persistent launch, "serving" the
request loop.  `roofline` prices the plain kernel against the FP32
matrix-core peak; `cpu_baseline` times the oracle's OpenMP FP32 restatement
of the reference's naive kernel on a bounded sample (and the FP64 oracle).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sgemm.cu_amd"))

PEAK_FP32_TFLOPS = 157.3      # MI355X dense FP32 (vector and matrix)
PEAK_BF16_TFLOPS = 2500.0     # MI355X dense BF16 MFMA
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-bf16x3", action="store_true")
    ap.add_argument("--no-error", action="store_true", help="skip the FP64 error report")
    ap.add_argument("--cpu-cols", type=int, default=0,
                    help="output columns of the CPU baseline sample (0 = all: the whole problem)")
    ap.add_argument("--large-size", type=int, default=16384,
                    help="also time both forms, and their FP64 error, at this size (0 = skip)")
    ap.add_argument("--no-grouped", action="store_true")
    ap.add_argument("--no-sweep", action="store_true", help="skip the compact size sweep")
    ap.add_argument("--no-vendor", action="store_true", help="skip the vendor BLAS line")
    ap.add_argument("--tune-iters", type=int, default=20,
                    help="timed launches per candidate when the grouped leg tunes its buckets")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--group-count", type=int, default=512)
    ap.add_argument("--serve-seconds", type=float, default=2.0,
                    help="arrival window per swept rate")
    ap.add_argument("--serve-chunk-gflop", type=float, default=0.0,
                    help="server chunk threshold in GFLOP (0 = library default, < 0 = off)")
    ap.add_argument("--serve-slots", type=int, default=3,
                    help="server slots (streams with a launch in flight)")
    ap.add_argument("--serve-rates", type=str, default="0.25,0.5,0.75,0.9,1.0",
                    help="arrival rates as fractions of the measured closed-loop capacity "
                         "(at 1.0 the queue does not settle: its p99 grows with the window)")
    ap.add_argument("--plain-only", action="store_true",
                    help="time only the plain FP32 kernel (profiling runs)")
    ap.add_argument("--x3-only", action="store_true",
                    help="time only the BF16x3 kernel (profiling runs)")
    ap.add_argument("--sweep", action="store_true",
                    help="size sweep instead of the single-size line: square 1024..16384 step "
                         "1024 plus (1000,1000,1000), (4097,3071,2049), (8192,8192,127); "
                         "alpha=1.5, beta=0.5; both forms")
    ap.add_argument("--shape", type=str, default="",
                    help="M,N,K: time only the default entry point of --form at this shape "
                         "(alpha=1.5, beta=0.5, ld = rows; profiling runs) and exit")
    ap.add_argument("--form", type=str, default="f32", choices=("f32", "bf16x3"),
                    help="form timed by --shape")
    ap.add_argument("--grouped-only", action="store_true",
                    help="run only the grouped leg (profiling runs) and exit")
    ap.add_argument("--serving-only", action="store_true",
                    help="run only the serving leg and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 1 GiB device-to-device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.plain_only:
        args.no_bf16x3 = args.no_error = args.no_cpu = args.no_grouped = args.no_serve = True
        args.no_sweep = args.no_vendor = True
    if args.x3_only:
        args.no_error = args.no_cpu = args.no_grouped = args.no_serve = True
        args.no_sweep = args.no_vendor = True
    return args


def sweep(sa, torch, warmup):
    """TFLOPS and fraction of peak per size for both forms (one JSON line)."""
    shapes = [(s, s, s) for s in range(1024, 16385, 1024)]
    shapes += [(1000, 1000, 1000), (4097, 3071, 2049), (8192, 8192, 127)]
    stream = torch.cuda.current_stream()
    out = []
    for (m, n, k) in shapes:
        A = torch.rand(m * k, device="cuda") * 2 - 1
        B = torch.rand(k * n, device="cuda") * 2 - 1
        C = torch.rand(m * n, device="cuda") * 2 - 1
        flops = 2.0 * m * n * k
        # About 20 TFLOP of timed work per size (at least 3 calls, at most
        # 500), after a quarter as many warm-up calls: 20 calls of a small
        # size finish before the engine clock has risen.
        it = max(3, min(500, int(2e13 / flops)))
        warm = max(warmup, it // 4)
        row = {"m": m, "n": n, "k": k}
        for form, name, fn, peak in ((sa.FORM_F32, "f32", sa.sgemm_ex, PEAK_FP32_TFLOPS),
                                     (sa.FORM_BF16X3, "bf16x3", sa.sgemm_bf16x3,
                                      PEAK_BF16_TFLOPS / 3)):
            ms = time_form(lambda: fn(m, n, k, 1.5, A, m, B, k, 0.5, C, m, stream), it, warm,
                           stream)
            tf = flops / (ms * 1e-3) / 1e12
            cfg = sa.launch_config(form, m, n, k, A, m, B, k)
            row[name] = {"ms": ms, "tflops": tf, "frac_of_peak": tf / peak, "config": cfg}
        print(f"sweep {m}x{n}x{k}: f32 {row['f32']['tflops']:.1f} TF, "
              f"bf16x3 {row['bf16x3']['tflops']:.1f} TF", file=sys.stderr, flush=True)
        out.append(row)
        del A, B, C
        torch.cuda.empty_cache()
    print(json.dumps({"metric": "sgemm size sweep", "unit": "TFLOPS",
                      "peaks": {"f32": PEAK_FP32_TFLOPS, "bf16x3_useful": PEAK_BF16_TFLOPS / 3},
                      "alpha": 1.5, "beta": 0.5, "data": "synthetic uniform [-1,1)",
                      "rows": out}), flush=True)


def sweep_compact(sa, torch, stream, target_s=0.08):
    """The size sweep in brief (both forms, TFLOPS and fraction of peak per
    size): each (size, form) is warmed up, then timed over about target_s
    seconds of back-to-back calls with HIP events on the launch stream."""
    shapes = [(s, s, s) for s in range(1024, 16385, 1024)]
    shapes += [(1000, 1000, 1000), (4097, 3071, 2049), (8192, 8192, 127)]
    rows = []
    for (m, n, k) in shapes:
        A = torch.rand(m * k, device="cuda") * 2 - 1
        B = torch.rand(k * n, device="cuda") * 2 - 1
        C = torch.rand(m * n, device="cuda") * 2 - 1
        flops = 2.0 * m * n * k
        row = {"m": m, "n": n, "k": k}
        for form, name, fn, peak in ((sa.FORM_F32, "f32", sa.sgemm_ex, PEAK_FP32_TFLOPS),
                                     (sa.FORM_BF16X3, "bf16x3", sa.sgemm_bf16x3,
                                      PEAK_BF16_TFLOPS / 3)):
            def call():
                fn(m, n, k, 1.5, A, m, B, k, 0.5, C, m, stream)
            probe = time_form(call, 2, 1, stream)  # ms per call, cold
            it = max(3, min(400, int(target_s * 1e3 / max(probe, 1e-3))))
            ms = time_form(call, it, max(2, it // 2), stream)
            tf = flops / (ms * 1e-3) / 1e12
            row[name] = {"tflops": tf, "frac_of_peak": tf / peak, "ms": ms, "calls": it,
                         "config": sa.launch_config(form, m, n, k, A, m, B, k)}
        rows.append(row)
        del A, B, C
    torch.cuda.empty_cache()
    return {"alpha": 1.5, "beta": 0.5, "data": "synthetic uniform [-1,1)",
            "peaks": {"f32": PEAK_FP32_TFLOPS, "bf16x3_useful": PEAK_BF16_TFLOPS / 3},
            "rows": rows}


def vendor_leg(torch, n, A, B, C0, steps, warmup, stream):
    """The vendor BLAS on the same inputs and problem (reference:
    benchmark.cu's CUBLAS=1 line): torch's FP32 addmm, which calls hipBLASLt
    or rocBLAS on ROCm, with each library that torch lets us prefer.  The
    column-major C = 1.5 A B + 0.5 C is the row-major C^T = 1.5 B^T A^T +
    0.5 C^T, which is what the row-major views below compute."""
    out = {}
    Av = A.view(n, n)   # row-major view of column-major A: A^T
    Bv = B.view(n, n)   # B^T
    flops = 2.0 * n ** 3
    prev = None
    try:
        prev = torch.backends.cuda.preferred_blas_library()
    except Exception:
        pass
    torch.backends.cuda.matmul.allow_tf32 = False
    for name, lib in (("hipblaslt", "cublaslt"), ("rocblas", "cublas")):
        try:
            got = torch.backends.cuda.preferred_blas_library(lib)
        except Exception as e:  # noqa: BLE001 - report and go on
            out[name] = {"error": str(e)}
            continue
        C = C0.clone()
        Cv = C.view(n, n)
        ms = time_form(lambda: Cv.addmm_(Bv, Av, beta=0.5, alpha=1.5), steps, warmup, stream)
        out[name] = {"library_set": str(got), "ms": ms,
                     "tflops": flops / (ms * 1e-3) / 1e12}
        del C, Cv
    if prev is not None:
        try:
            torch.backends.cuda.preferred_blas_library(prev)
        except Exception:
            pass
    torch.cuda.empty_cache()
    return out


def copy_calibration(steps):
    """A copy kernel of known size: 1 GiB read + 1 GiB written per step, the
    yardstick for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = 1 << 28
    src = torch.rand(n, device="cuda")
    dst = torch.empty_like(src)
    ms = time_form(lambda: dst.copy_(src), steps, 2, torch.cuda.current_stream())
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "ms": ms,
                      "GBps": 2 * 4 * n / (ms * 1e-3) / 1e9}), flush=True)


class SclkSampler:
    """Engine clock during a timed region: a thread reads the current level of
    the amdgpu sysfs DPM table (pp_dpm_sclk, the line marked '*') every 2 ms.
    Reports nothing when the file is not readable."""

    def __init__(self):
        import glob
        paths = sorted(glob.glob("/sys/class/drm/card*/device/pp_dpm_sclk"))
        self.path = paths[0] if paths else None
        self.samples = []
        self._stop = None
        self._thread = None

    def _read(self):
        with open(self.path) as f:
            for line in f:
                if line.rstrip().endswith("*"):
                    return float(line.split(":")[1].strip().rstrip("*").strip().lower()
                                 .replace("mhz", ""))
        return None

    def __enter__(self):
        import threading
        if self.path is None:
            return self
        try:
            self._read()
        except (OSError, ValueError, IndexError):
            self.path = None
            return self
        self._stop = threading.Event()

        def run():
            while not self._stop.wait(0.002):
                try:
                    v = self._read()
                except (OSError, ValueError, IndexError):
                    v = None
                if v is not None:
                    self.samples.append(v)

        self._thread = threading.Thread(target=run, daemon=True)
        self._thread.start()
        return self

    def __exit__(self, *exc):
        if self._thread is not None:
            self._stop.set()
            self._thread.join()
        return False

    def summary(self):
        if not self.samples:
            return None
        v = sorted(self.samples)
        return {"mean_mhz": sum(v) / len(v), "p50_mhz": v[len(v) // 2], "min_mhz": v[0],
                "max_mhz": v[-1], "samples": len(v)}


def time_form(fn, steps, warmup, stream):
    """Time `steps` back-to-back calls with HIP events on the launch stream."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        fn()
    stop.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps  # ms per call


def settle(fn, stream, warmup, seconds=0.3):
    """At least `warmup` calls and about `seconds` of back-to-back calls, so
    the engine clock has settled before a short launch is timed (the first
    few milliseconds after a phase of small launches run slower)."""
    probe = time_form(fn, 2, 1, stream)
    time_form(fn, max(warmup, int(seconds * 1e3 / max(probe, 1e-3))), 0, stream)


def per_call(fn, calls, stream, flush=None):
    """Each call timed on its own (one HIP event pair per call): the median,
    and the mean of the second half as the reference's harness reports it
    (benchmark.cu:141-183).  `flush`, when given, is run between calls,
    outside the timed pairs, to empty the L2 and Infinity caches as the
    reference's harness does (common/helper_cuda.h:70-83)."""
    import numpy as np
    import torch
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(calls)]
    for a, b in evs:
        if flush is not None:
            flush()
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    t = [a.elapsed_time(b) for a, b in evs]
    return {"calls": calls, "p50_ms": float(np.median(t)),
            "mean_second_half_ms": float(np.mean(t[calls // 2:])), "cache_flush": flush is not None}


def error_vs_fp64(C, ref64, scale64):
    """Componentwise error |C - ref| / (|alpha||A||B| + |beta||C0|) and the
    plain relative error over ref != 0 (max and mean of both)."""
    import torch
    d = (C.double() - ref64).abs()
    comp = d / scale64
    nz = ref64 != 0
    rel = d[nz] / ref64[nz].abs()
    return {"max_comp": float(comp.max()), "mean_comp": float(comp.mean()),
            "max_rel": float(rel.max()), "mean_rel": float(rel.mean())}


def host_fp64_entries(A, B, n, ii, jj, chunk=512):
    """FP64 reference and scale |A||B| of entries (ii, jj) of the column-major
    n x n product of host float32 operands: float64 dot products (numpy), the
    computation the FP64 host oracle restates for the tests."""
    import numpy as np
    Am = A.reshape(n, n)   # Am[p, i] = A(i, p)
    Bm = B.reshape(n, n)   # Bm[j, p] = B(p, j)
    ref = np.empty(len(ii))
    scale = np.empty(len(ii))
    for s in range(0, len(ii), chunk):
        a = Am[:, ii[s:s + chunk]].T.astype(np.float64)
        b = Bm[jj[s:s + chunk], :].astype(np.float64)
        ref[s:s + chunk] = np.einsum("ij,ij->i", a, b)
        scale[s:s + chunk] = np.einsum("ij,ij->i", np.abs(a), np.abs(b))
    return ref, scale


def wide_error(sa, torch, n, with_x3, stream, samples=8192):
    """Wide-range inputs (sign * 2^u, u in [-8, 8)): the componentwise and
    relative error of each form at n^3 over the whole matrix against an FP64
    torch matmul on the GPU, and over `samples` fixed entries against FP64
    dot products computed on the host from the same (downloaded) operands."""
    import numpy as np
    dev = "cuda"
    g = torch.Generator(device=dev)
    g.manual_seed(1)

    def wide():
        u = torch.rand(n * n, device=dev, generator=g, dtype=torch.float64) * 16 - 8
        s = torch.where(torch.rand(n * n, device=dev, generator=g) < 0.5, -1.0, 1.0)
        return (s.double() * torch.exp2(u)).float()
    Aw, Bw = wide(), wide()
    rng = np.random.default_rng(n)
    ii = rng.integers(0, n, samples)
    jj = rng.integers(0, n, samples)
    href, hscale = host_fp64_entries(Aw.cpu().numpy(), Bw.cpu().numpy(), n, ii, jj)
    idx = torch.from_numpy(ii + jj * n).to(dev)
    Am = Aw.view(n, n).t().double()   # column-major -> logical (m, k)
    Bm = Bw.view(n, n).t().double()
    ref = (Am @ Bm).t().contiguous().view(-1)
    scale = (Am.abs() @ Bm.abs()).t().contiguous().view(-1)
    del Am, Bm
    Cw = torch.zeros(n * n, device=dev)
    out = {}
    forms = [("f32", sa.sgemm_ex)] + ([("bf16x3", sa.sgemm_bf16x3)] if with_x3 else [])
    for name, fn in forms:
        fn(n, n, n, 1.0, Aw, n, Bw, n, 0.0, Cw, n, stream)
        e = error_vs_fp64(Cw, ref, scale)
        got = Cw[idx].double().cpu().numpy()
        comp = np.abs(got - href) / hscale
        e["sampled_host_fp64"] = {"samples": samples, "max_comp": float(comp.max()),
                                  "mean_comp": float(comp.mean())}
        out[name] = e
    del ref, scale, Aw, Bw, Cw
    torch.cuda.empty_cache()
    return out


def forms_at(sa, torch, n, steps, warmup, stream, with_error):
    """Both forms at n^3 (uniform [-1, 1) inputs, alpha=1.5, beta=0.5), and
    their wide-range error against FP64."""
    A = torch.rand(n * n, device="cuda") * 2 - 1
    B = torch.rand(n * n, device="cuda") * 2 - 1
    C = torch.rand(n * n, device="cuda") * 2 - 1
    flops = 2.0 * n ** 3
    ms = time_form(lambda: sa.sgemm_ex(n, n, n, 1.5, A, n, B, n, 0.5, C, n, stream),
                   steps, warmup, stream)
    ms3 = time_form(lambda: sa.sgemm_bf16x3(n, n, n, 1.5, A, n, B, n, 0.5, C, n, stream),
                    steps, warmup, stream)
    del A, B, C
    torch.cuda.empty_cache()
    out = {"size": n, "steps": steps,
           "f32": {"ms": ms, "tflops": flops / (ms * 1e-3) / 1e12},
           "bf16x3": {"ms": ms3, "tflops": flops / (ms3 * 1e-3) / 1e12},
           "speedup_bf16x3_vs_f32": ms / ms3}
    if with_error:
        for name, e in wide_error(sa, torch, n, True, stream).items():
            out[name]["error_wide"] = e
    return out


def cpu_baseline(size: int, cols: int):
    """Oracle FP32 restatement of kernel_basic (OpenMP) on `cols` output columns
    of the size^3 problem: a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as orc
    m, k = size, size
    cols = cols if cols > 0 else size
    A = orc.fill("uniform", m * k, 1)
    B = orc.fill("uniform", k * cols, 2)
    C = orc.fill("uniform", m * cols, 3)
    t0 = time.perf_counter()
    orc.sgemm_basic_cm(m, cols, k, 1.5, A, m, B, k, 0.5, C, m)
    dt = time.perf_counter() - t0
    flops = 2.0 * m * cols * k
    # the 1024^3 case as well (median of 5), as a second reference point
    s = 1024
    A1, B1 = orc.fill("uniform", s * s, 1), orc.fill("uniform", s * s, 2)
    times = []
    for _ in range(5):
        C1 = orc.fill("uniform", s * s, 3)
        t1 = time.perf_counter()
        orc.sgemm_basic_cm(s, s, s, 1.5, A1, s, B1, s, 0.5, C1, s)
        times.append(time.perf_counter() - t1)
    t1024 = float(np.median(times))
    # the FP64 host oracle (the ground truth of the tests) on a bounded sample
    cols64 = min(256, cols)
    t2 = time.perf_counter()
    orc.dgemm(m, cols64, k, 1.5, A, m, B, k, 0.5, C, m, with_scale=False)
    dt64 = time.perf_counter() - t2
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": flops / dt / 1e12, "unit": "TFLOPS", "cores": orc.max_threads(),
            "kind": "port",
            "sample": f"{cols} of {size} output columns of the {size}^3 problem "
                      f"(M={m}, N={cols}, K={k}), alpha=1.5 beta=0.5, {dt:.1f} s",
            "tflops_1024": 2.0 * s ** 3 / t1024 / 1e12, "ms_1024_median_of_5": t1024 * 1e3,
            "fp64_oracle": {"tflops": 2.0 * m * cols64 * k / dt64 / 1e12, "seconds": dt64,
                            "sample": f"M={m}, N={cols64}, K={k}, alpha=1.5 beta=0.5"},
            "host_cpus_visible": os.cpu_count(), "cpu_model": model}


def grouped_leg(sa, torch, form, count, steps, warmup, stream, seed=0, tune_iters=20):
    """512 independent problems, M, N, K log-uniform integers in [32, 2048],
    U[-1, 1), alpha=1, beta=0: one persistent launch replayed from a hipGraph,
    against the same problems as back-to-back single launches of the default
    entry point (also graph-captured, so no host overhead is timed)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    shapes = np.floor(np.exp(rng.uniform(np.log(32), np.log(2049), size=(count, 3)))).astype(int)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed + 1)
    mats, probs = [], []
    flops = 0.0
    for m, n, k in shapes.tolist():
        A = torch.rand(m * k, device="cuda", generator=g) * 2 - 1
        B = torch.rand(k * n, device="cuda", generator=g) * 2 - 1
        C = torch.zeros(m * n, device="cuda")
        mats.append((A, B, C))
        probs.append(sa.problem(m, n, k, 1.0, A, m, B, k, 0.0, C, m))
        flops += 2.0 * m * n * k
    # autotune every size bucket present (grouped-capable configs only),
    # after about 0.3 s of 2048^3 launches so that the first candidates are
    # not timed on a clock that is still rising
    single = sa.sgemm_ex if form == sa.FORM_F32 else sa.sgemm_bf16x3
    W = torch.rand(2048 * 2048, device="cuda")
    settle(lambda: single(2048, 2048, 2048, 1.0, W, 2048, W, 2048, 0.0, W, 2048, stream), stream,
           warmup)
    del W
    t0 = time.perf_counter()
    tuned = {}
    for m, n, k in shapes.tolist():
        key = tuple(int(v - 1).bit_length() for v in (m, n, k))
        if key not in tuned:
            tuned[key] = sa.autotune_bucket(form, m, n, k, iters=tune_iters, grouped=True,
                                            stream=stream)[0]
    tune_s = time.perf_counter() - t0
    torch.cuda.synchronize()
    gs = torch.cuda.Stream()  # stream capture needs a stream other than the null stream
    plan = sa.GroupedPlan(count, form)
    plan.set(probs, gs)
    plan.capture(gs)
    settle(lambda: plan.replay(gs), gs, warmup)
    ms_group = time_form(lambda: plan.replay(gs), steps, 0, gs)
    # the same problems as single launches of the library's default entry
    # point (each with its own single-launch config), captured in one graph on
    # a stream warmed up first (so split-K workspaces exist before capture)
    loop_graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(stream)
    with torch.cuda.stream(side):
        for (m, n, k), (A, B, C) in zip(shapes.tolist(), mats):  # warm-up
            single(m, n, k, 1.0, A, m, B, k, 0.0, C, m, side)
    side.synchronize()
    with torch.cuda.graph(loop_graph, stream=side):
        for (m, n, k), (A, B, C) in zip(shapes.tolist(), mats):
            single(m, n, k, 1.0, A, m, B, k, 0.0, C, m, side)
    settle(loop_graph.replay, stream, warmup)
    ms_loop = time_form(loop_graph.replay, steps, 0, stream)
    info = plan.info()
    plan.destroy()
    return {"problems": count, "flops": flops, "buckets_tuned": len(tuned),
            "tune_s": tune_s, "tune_iters": tune_iters, "ms_grouped": ms_group,
            "tflops_grouped": flops / (ms_group * 1e-3) / 1e12,
            "ms_single_loop": ms_loop, "tflops_single_loop": flops / (ms_loop * 1e-3) / 1e12,
            "speedup_vs_loop": ms_loop / ms_group, "tiles": info["total_tiles"],
            "grid": info["grid"]}


def serving_leg(sa, torch, rates, window_s, chunk_gflop=0.0, seed=0, slots=3):
    """Poisson arrivals of BF16x3 requests, sizes 70% [128,512], 25%
    [1024,2048], 5% [4096,8192] (each of M, N, K uniform in the bucket);
    operands resident in HBM, 64 GB budget.  Each request reads its A and B
    from a random place in 2 GiB pools (eight 8192^2 panels each: a random
    panel, a random row offset that is a multiple of 4 and a random column
    offset), so requests do not share operand lines in the caches."""
    import numpy as np
    rng = np.random.default_rng(seed)
    side = 8192
    panels = 8
    Apool = torch.rand(panels * side * side, device="cuda") * 2 - 1
    Bpool = torch.rand(panels * side * side, device="cuda") * 2 - 1
    arena = torch.zeros(2 << 30, device="cuda")   # 8 GiB ring for results
    ring = [0]

    def operand(pool, rows, cols):
        """Address of a rows x cols block (leading dimension `side`)."""
        p = int(rng.integers(0, panels))
        r0 = 4 * int(rng.integers(0, (side - rows) // 4 + 1))
        c0 = int(rng.integers(0, side - cols + 1))
        return pool.data_ptr() + 4 * (p * side * side + c0 * side + r0)

    buckets = ((0.70, 128, 512), (0.25, 1024, 2048), (0.05, 4096, 8192))

    def draw():
        r = rng.random()
        lo, hi = (128, 512) if r < 0.70 else (1024, 2048) if r < 0.95 else (4096, 8192)
        return tuple(int(v) for v in rng.integers(lo, hi + 1, 3))
    # E[2mnk] of the mixture (m, n, k independent and uniform in the bucket):
    # 2.5e10 flops per request, 93 % of it in the 5 % bucket
    mean_flops = sum(p * 2.0 * ((lo + hi) / 2.0) ** 3 for p, lo, hi in buckets)

    def c_slice(m, n):
        if ring[0] + m * n > arena.numel():
            ring[0] = 0
        p = arena.data_ptr() + 4 * ring[0]
        ring[0] += (m * n + 63) // 64 * 64
        return p

    srv = sa.Server(64 << 30, max_batch=32, num_slots=slots, form=sa.FORM_BF16X3,
                    chunk_gflop=chunk_gflop)

    def submit(m, n, k):
        return srv.submit(m, n, k, 1.0, operand(Apool, m, k), side, operand(Bpool, k, n), side,
                          0.0, c_slice(m, n), m)

    # closed-loop capacity: a burst of requests, all queued at once.  Its
    # TFLOP/s is the capacity; in requests per second it is that over the
    # mixture's mean flops per request -- a burst's own request rate depends
    # on how many large requests it happened to draw (a 600-request burst
    # read 9400 req/s, 30 % above the mean-flops rate, so "0.75x" was
    # near saturation).
    for _ in range(50):
        submit(*draw())
    srv.drain(120)
    srv.reset_stats()
    burst = [draw() for _ in range(1500)]
    for s in burst:
        submit(*s)
    srv.drain(300)
    cap = srv.stats()
    capacity = cap["tflops"] * 1e12 / mean_flops
    out = {"capacity_req_s": capacity, "capacity_tflops": cap["tflops"],
           "capacity_req_s_burst": cap["completed"] / cap["window_s"],
           "mean_flops_per_request": mean_flops, "budget_gb": 64,
           "operands": "random blocks of 2 x 2 GiB resident pools (8 panels of 8192^2 each)",
           "max_batch": 32, "slots": slots, "chunk_gflop": chunk_gflop, "rates": []}
    for frac in rates:
        lam = frac * capacity
        n_req = max(50, int(lam * window_s))
        gaps = rng.exponential(1.0 / lam, size=n_req)
        sizes = [draw() for _ in range(n_req)]
        srv.reset_stats()
        t_start = time.perf_counter()
        t_next = t_start
        ids = []
        for gap, s in zip(gaps, sizes):
            t_next += gap
            while time.perf_counter() < t_next:
                pass
            ids.append(submit(*s))
        srv.drain(300)
        st = srv.stats()
        # per size class: the 4096..8192 bucket ("large") against the rest
        lat = {"small": [], "large": []}
        for rid, (m, n, k) in zip(ids, sizes):
            lat["large" if min(m, n, k) >= 4096 else "small"].append(srv.latency(rid)[0])

        def pct(v, p):
            if not v:
                return None
            v = sorted(v)
            return v[min(len(v), max(1, int(np.ceil(p / 100 * len(v))))) - 1]
        out["rates"].append({"fraction_of_capacity": frac, "rate_req_s": lam,
                             "requests": st["measured"], "p50_ms": st["p50_ms"],
                             "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
                             "mean_queue_ms": st["mean_queue_ms"], "tflops": st["tflops"],
                             "mean_batch": st["mean_batch"],
                             "panel_launches": st["panel_launches"],
                             "p50_small_ms": pct(lat["small"], 50),
                             "p99_small_ms": pct(lat["small"], 99),
                             "p50_large_ms": pct(lat["large"], 50),
                             "p99_large_ms": pct(lat["large"], 99),
                             "peak_gb": st["peak_bytes"] / 2 ** 30})
    srv.destroy()
    del Apool, Bpool, arena
    torch.cuda.empty_cache()
    return out


def error_target(forms, large):
    """The split-precision target, per size: BF16x3 at least 2x the plain
    form's TFLOPS with a max componentwise error no larger than the plain
    form's, on the whole matrix (FP64 torch matmul) and on the sampled
    entries (host FP64)."""
    out = {}
    for label, f in (("n8192", forms), ("n16384", large)):
        if not f or "bf16x3" not in f or "error_wide" not in f.get("bf16x3", {}):
            continue
        e32, e3 = f["f32"]["error_wide"], f["bf16x3"]["error_wide"]
        speed = f["bf16x3"]["tflops"] / f["f32"]["tflops"]
        full = e3["max_comp"] <= e32["max_comp"]
        samp = (e3["sampled_host_fp64"]["max_comp"] <=
                e32["sampled_host_fp64"]["max_comp"])
        out[label] = {"speedup": speed, "max_full_matrix_ok": full, "max_sampled_ok": samp,
                      "mean_ratio_full_matrix": e3["mean_comp"] / e32["mean_comp"],
                      "met": bool(speed >= 2.0 and full and samp)}
    return out


def load_traffic(size: int):
    path = os.path.join(ROOT, "profiles", f"traffic_sgemm_f32_{size}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is not implemented)", file=sys.stderr)
        sys.exit(2)
    import torch
    import sgemm_amd as sa
    if not torch.cuda.is_available():
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration:
        copy_calibration(args.steps)
        return
    if args.sweep:
        sa.lib()
        sweep(sa, torch, args.warmup)
        return
    if args.shape:
        sa.lib()
        m, n, k = (int(v) for v in args.shape.split(","))
        form = sa.FORM_F32 if args.form == "f32" else sa.FORM_BF16X3
        fn = sa.sgemm_ex if form == sa.FORM_F32 else sa.sgemm_bf16x3
        A = torch.rand(m * k, device="cuda") * 2 - 1
        B = torch.rand(k * n, device="cuda") * 2 - 1
        C = torch.rand(m * n, device="cuda") * 2 - 1
        stream = torch.cuda.current_stream()

        def call():
            fn(m, n, k, 1.5, A, m, B, k, 0.5, C, m, stream)
        # Warm up for about 0.3 s so the engine clock has settled (a few
        # calls of a short problem end before it rises), then time `steps`.
        settle(call, stream, args.warmup)
        ms = time_form(call, args.steps, 0, stream)
        print(json.dumps({"shape": [m, n, k], "form": args.form, "ms": ms,
                          "tflops": 2.0 * m * n * k / (ms * 1e-3) / 1e12,
                          "config": sa.launch_config(form, m, n, k, A, m, B, k)}), flush=True)
        return
    if args.grouped_only:
        sa.lib()
        stream = torch.cuda.current_stream()
        out = {}
        for form, name in ((sa.FORM_F32, "f32"), (sa.FORM_BF16X3, "bf16x3")):
            out[name] = grouped_leg(sa, torch, form, args.group_count, args.steps, args.warmup,
                                    stream, tune_iters=args.tune_iters)
        print(json.dumps({"grouped": out}), flush=True)
        return
    if args.serving_only:
        sa.lib()
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        print(json.dumps({"serving": serving_leg(sa, torch, rates, args.serve_seconds,
                                                 args.serve_chunk_gflop,
                                                 slots=args.serve_slots)}), flush=True)
        return
    sa.lib()
    torch.manual_seed(0)
    n = args.size
    dev = "cuda"
    A = torch.rand(n * n, device=dev, dtype=torch.float32) * 2 - 1
    B = torch.rand(n * n, device=dev, dtype=torch.float32) * 2 - 1
    C0 = torch.rand(n * n, device=dev, dtype=torch.float32) * 2 - 1
    C = C0.clone()
    alpha, beta = 1.5, 0.5
    stream = torch.cuda.current_stream()
    flops = 2.0 * n * n * n

    def plain():
        sa.sgemm_ex(n, n, n, alpha, A, n, B, n, beta, C, n, stream)

    print(f"bench: plain FP32 {n}^3, warmup {args.warmup}, steps {args.steps}", file=sys.stderr,
          flush=True)
    with SclkSampler() as clk:
        ms = time_form(plain, 1 if args.x3_only else args.steps,
                       0 if args.x3_only else args.warmup, stream)
    tflops = flops / (ms * 1e-3) / 1e12
    forms = {"f32": {"tflops": tflops, "ms": ms,
                     "config": sa.launch_config(sa.FORM_F32, n, n, n, A, n, B, n),
                     "sclk_sysfs": clk.summary()}}
    # 512 MiB written between calls: more than the 8 x 4 MB L2 plus the 256 MB
    # Infinity Cache, so every flushed call starts with its operands in HBM
    flush_buf = torch.empty(128 << 20, device=dev, dtype=torch.float32)

    def flush():
        flush_buf.fill_(0.0)
    if not args.x3_only:
        forms["f32"]["per_call"] = per_call(plain, min(args.steps, 20), stream)
        forms["f32"]["per_call_flushed"] = per_call(plain, min(args.steps, 10), stream, flush)

    if not args.no_bf16x3 and sa.lib().sgemm_config_count(sa.FORM_BF16X3) > 0:
        def split():
            sa.sgemm_bf16x3(n, n, n, alpha, A, n, B, n, beta, C, n, stream)
        print("bench: BF16x3", file=sys.stderr, flush=True)
        with SclkSampler() as clk3:
            ms3 = time_form(split, args.steps, args.warmup, stream)
        t3 = flops / (ms3 * 1e-3) / 1e12
        forms["bf16x3"] = {"tflops": t3, "ms": ms3, "speedup_vs_f32": t3 / tflops,
                           "config": sa.launch_config(sa.FORM_BF16X3, n, n, n, A, n, B, n),
                           "mfma_tflops_issued": 3 * t3, "sclk_sysfs": clk3.summary(),
                           "per_call": per_call(split, min(args.steps, 20), stream),
                           "per_call_flushed": per_call(split, min(args.steps, 10), stream,
                                                        flush)}

    vendor = None
    if not args.no_vendor:
        print("bench: vendor BLAS", file=sys.stderr, flush=True)
        vendor = vendor_leg(torch, n, A, B, C0, args.steps, args.warmup, stream)
        for v in vendor.values():
            if "tflops" in v:
                v["ours_over_vendor"] = tflops / v["tflops"]

    if not args.no_error:
        print("bench: error vs FP64", file=sys.stderr, flush=True)
        errs = wide_error(sa, torch, n, "bf16x3" in forms, stream)
        for name, e in errs.items():
            forms[name]["error_wide"] = e

    large = None
    if args.large_size > 0 and not args.no_bf16x3 and not args.plain_only and not args.x3_only:
        print(f"bench: both forms at {args.large_size}^3", file=sys.stderr, flush=True)
        large = forms_at(sa, torch, args.large_size, min(args.steps, 5), min(args.warmup, 2),
                         stream, not args.no_error)

    sweep_rows = None
    if not args.no_sweep:
        print("bench: size sweep", file=sys.stderr, flush=True)
        sweep_rows = sweep_compact(sa, torch, stream)

    grouped = None
    if not args.no_grouped:
        grouped = {}
        for form, name in ((sa.FORM_F32, "f32"), (sa.FORM_BF16X3, "bf16x3")):
            print(f"bench: grouped {name}", file=sys.stderr, flush=True)
            grouped[name] = grouped_leg(sa, torch, form, args.group_count, args.steps,
                                        args.warmup, stream, tune_iters=args.tune_iters)

    serving = None
    if not args.no_serve:
        print("bench: serving", file=sys.stderr, flush=True)
        rates = [float(x) for x in args.serve_rates.split(",") if x]
        serving = serving_leg(sa, torch, rates, args.serve_seconds, args.serve_chunk_gflop,
                              slots=args.serve_slots)

    achieved = tflops
    roofline = {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                "traffic": load_traffic(n)}

    cpu = None
    if not args.no_cpu:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(n, args.cpu_cols)

    line = {
        "metric": "FP32 GEMM TFLOPS @8192^3 plain vs BF16x3-MFMA; p50/p99 grouped-request latency",
        "value": tflops,
        "unit": "TFLOPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (uniform [-1,1), torch seed 0, resident in HBM)",
        "config": {"workload": f"sgemm column-major M=N=K={n}, lda=ldb=ldc={n}, alpha=1.5, "
                               f"beta=0.5", "m": n, "n": n, "k": n},
        "forms": forms,
        "forms_large": large,
        "error_target": error_target(forms, large),
        "vendor": vendor,
        "sweep": sweep_rows,
        "grouped": grouped,
        "serving": serving,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
