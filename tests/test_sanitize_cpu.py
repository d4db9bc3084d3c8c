"""The serving scheduler under ThreadSanitizer and AddressSanitizer +
UndefinedBehaviorSanitizer, on the CPU: tests/sanitize/Makefile builds the
library with instrumented host code and the multi-threaded stress driver
serve_stress.cpp, which runs on the mock executor (no GPU).  A sanitizer
report ends the run with a non-zero status."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")


def _build(target):
    if shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("hipcc is not installed")
    subprocess.run(["make", "-C", SAN, "-j8", target], check=True, capture_output=True,
                   timeout=900)
    return os.path.join(SAN, "_build", f"serve_stress_{target}")


@pytest.mark.parametrize("target,env", [
    ("tsan", {"TSAN_OPTIONS": "halt_on_error=1 exitcode=66"}),
    ("asan", {"ASAN_OPTIONS": "detect_leaks=1 halt_on_error=1",
              "UBSAN_OPTIONS": "halt_on_error=1 print_stacktrace=1"}),
])
def test_serving_under_sanitizer(target, env):
    exe = _build(target)
    r = subprocess.run([exe], env={**os.environ, **env}, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "ok" in r.stdout
