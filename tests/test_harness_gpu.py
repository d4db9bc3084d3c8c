"""The C++ harnesses that replace the reference's test.cu and benchmark.cu,
run as their own processes (they link the library and /opt/rocm's HIP
runtime, no torch).

* sgemm_test: every square size in a range through sgemm against the naive
  sgemm_basic by the reference's acceptance rule (relative error > 1e-4 or
  NaN/Inf fails, common/helper_matrix.h:51-86), summary written to
  <savedir>/sgemm.cu.txt as test.cu:108-118 does; non-zero exit on failure.
* sgemm_benchmark: the reference's sweep protocol (benchmark.cu:141-183) and
  its "<size> <GFLOP/s>" output file (benchmark.cu:199-206), for the library
  and for rocBLAS (the CUBLAS=1 line of benchmark.cu:150-164)."""
import os
import re
import subprocess

import pytest

from gpu_util import log

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "sgemm.cu_amd", "bin")


def _exe(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built (make -C sgemm.cu_amd)")
    return path


@pytest.mark.parametrize("form", ["f32", "bf16x3"])
def test_correctness_harness_sizes_2_to_600(tmp_path, form):
    log(f"sgemm_test --form={form} --mmin=2 --mmax=600")
    r = subprocess.run([_exe("sgemm_test"), "--mmin=2", "--mmax=600", "--mstep=1",
                        f"--form={form}", f"--savedir={tmp_path}"],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    text = (tmp_path / "sgemm.cu.txt").read_text()
    m = re.search(r"PASSED: (\d+) / (\d+)", text)
    assert m and m.group(1) == m.group(2) == "599", text[-500:]
    assert "FAILED: 0" in text


@pytest.mark.parametrize("form", ["f32", "bf16x3", "rocblas"])
def test_benchmark_harness_file_format(tmp_path, form):
    log(f"sgemm_benchmark --form={form} --npts=3 --mstep=1024")
    r = subprocess.run([_exe("sgemm_benchmark"), "--npts=3", "--mstep=1024", "--mmin=1024",
                        "--wniter=5", "--wmsize=1024", f"--form={form}", f"--savedir={tmp_path}"],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    name = "rocBLAS.txt" if form == "rocblas" else "sgemm.cu.txt"
    lines = (tmp_path / name).read_text().splitlines()
    assert len(lines) == 3
    for line, size in zip(lines, (1024, 2048, 3072)):
        m = re.fullmatch(r"(\d+) (\d+)", line)
        assert m and int(m.group(1)) == size and int(m.group(2)) > 1000, line
