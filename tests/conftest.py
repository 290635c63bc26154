"""Test configuration: path set-up and the ``gpu`` marker.

CPU tests (``-m "not gpu"``) check the oracle against the reference's golden
fixtures and mathematical invariants, the host logic and the C ABI library's
exports.  GPU tests (``-m gpu``) call the HIP step through the C ABI and
compare it with the oracle.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("taxila-lbm_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP step)")
