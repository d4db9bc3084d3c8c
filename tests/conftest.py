"""pytest configuration: the `gpu` marker and import paths.

`-m "not gpu"` runs on any host (oracle, fixtures, host logic, ABI surface);
`-m gpu` needs an MI355X and calls the library through its C ABI.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "sgemm.cu_amd"),
          os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (gfx950); calls the HIP library")
