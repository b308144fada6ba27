"""pytest configuration: the `gpu` marker and import paths.

CPU tests (`-m "not gpu"`) pin the oracle and check the host logic and the
C-ABI library's exported symbols; GPU tests call the HIP library through its
C ABI and compare with the oracle.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "marie-tools_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
