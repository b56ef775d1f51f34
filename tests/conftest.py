"""Shared test configuration.

Markers: `gpu` tests need an AMD GPU and the built libksj_gpu.so; everything else
runs on the CPU (oracle checks, host logic, ABI symbol checks).
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: requires an AMD Instinct GPU and the built HIP libraries")
