#!/usr/bin/env python3
"""Benchmark of the column-major SGEMM hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 8192]

One step = one C := alpha*A*B + beta*C call at M=N=K=size (default 8192, the
shape the headline metric is quoted on) with alpha=1.5, beta=0.5 and entries
uniform in [-1, 1) (synthetic, seeded; resident in HBM before timing starts).
The flop convention is the reference's 2*M*N*K (benchmark.cu:138).

Prints ONE JSON line.  `value` is the plain FP32 form's TFLOPS; the BF16x3
form, its speedup and the error of both forms against an FP64 matmul of the
same inputs ride along under "forms".  `roofline` prices the plain kernel
against the FP32 matrix-core peak; `cpu_baseline` times the oracle's OpenMP
FP32 restatement of the reference's naive kernel on a bounded sample.
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
    ap.add_argument("--cpu-cols", type=int, default=48,
                    help="output columns of the CPU baseline sample")
    return ap.parse_args()


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


def cpu_baseline(size: int, cols: int):
    """Oracle FP32 restatement of kernel_basic (OpenMP) on `cols` output columns
    of the size^3 problem: a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as orc
    m, k = size, size
    A = orc.fill("uniform", m * k, 1)
    B = orc.fill("uniform", k * cols, 2)
    C = orc.fill("uniform", m * cols, 3)
    t0 = time.perf_counter()
    orc.sgemm_basic_cm(m, cols, k, 1.5, A, m, B, k, 0.5, C, m)
    dt = time.perf_counter() - t0
    flops = 2.0 * m * cols * k
    return {"value": flops / dt / 1e12, "unit": "TFLOPS", "cores": orc.max_threads(),
            "kind": "port",
            "sample": f"{cols} output columns of the {size}^3 problem (M={m}, N={cols}, K={k}), "
                      f"alpha=1.5 beta=0.5, {dt:.1f} s"}


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
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch
    import sgemm_amd as sa
    if not torch.cuda.is_available():
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
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
    ms = time_form(plain, args.steps, args.warmup, stream)
    tflops = flops / (ms * 1e-3) / 1e12
    forms = {"f32": {"tflops": tflops, "ms": ms}}

    if not args.no_bf16x3 and sa.lib().sgemm_config_count(sa.FORM_BF16X3) > 0:
        def split():
            sa.sgemm_bf16x3(n, n, n, alpha, A, n, B, n, beta, C, n, stream)
        print("bench: BF16x3", file=sys.stderr, flush=True)
        ms3 = time_form(split, args.steps, args.warmup, stream)
        t3 = flops / (ms3 * 1e-3) / 1e12
        forms["bf16x3"] = {"tflops": t3, "ms": ms3, "speedup_vs_f32": t3 / tflops,
                           "mfma_tflops_issued": 3 * t3}

    if not args.no_error:
        # Wide-range inputs (sign * 2^u, u in [-8, 8)), FP64 torch matmul as truth.
        print("bench: error vs FP64", file=sys.stderr, flush=True)
        g = torch.Generator(device=dev)
        g.manual_seed(1)
        def wide():
            u = torch.rand(n * n, device=dev, generator=g, dtype=torch.float64) * 16 - 8
            s = torch.where(torch.rand(n * n, device=dev, generator=g) < 0.5, -1.0, 1.0)
            return (s.double() * torch.exp2(u)).float()
        Aw, Bw = wide(), wide()
        Am = Aw.view(n, n).t().double()   # column-major -> logical (m, k)
        Bm = Bw.view(n, n).t().double()
        ref = (Am @ Bm).t().contiguous().view(-1)
        scale = (Am.abs() @ Bm.abs()).t().contiguous().view(-1)
        del Am, Bm
        Cw = torch.zeros_like(C)
        sa.sgemm_ex(n, n, n, 1.0, Aw, n, Bw, n, 0.0, Cw, n, stream)
        forms["f32"]["error_wide"] = error_vs_fp64(Cw, ref, scale)
        if "bf16x3" in forms:
            sa.sgemm_bf16x3(n, n, n, 1.0, Aw, n, Bw, n, 0.0, Cw, n, stream)
            forms["bf16x3"]["error_wide"] = error_vs_fp64(Cw, ref, scale)
        del ref, scale, Aw, Bw, Cw
        torch.cuda.empty_cache()

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
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
