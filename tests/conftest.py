"""Test configuration: import paths for the binding, the generators and the
CPU oracle, and the `gpu` marker (tests that need an MI355X)."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("info-centric_ge_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (MI355X) and the built libfelog_hip.so")
