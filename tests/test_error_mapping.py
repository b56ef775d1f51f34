"""Device out-of-memory reaches callers as KSJ_GPU_RESOURCE_EXHAUSTED (so the Provider
reports KSJ_STATUS_RESOURCE_EXHAUSTED and the server status 4), other HIP errors as
KSJ_GPU_DEVICE_ERROR and violated preconditions as KSJ_GPU_INVALID_ARGUMENT.  The
library's host-side mapping is compiled with g++ into a probe; no GPU is needed."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_hip_errors_map_to_the_abi_statuses(tmp_path):
    if shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include/hip"):
        pytest.skip("needs g++ and the HIP headers")
    exe = tmp_path / "probe"
    subprocess.run(["g++", "-std=c++20", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-I" + os.path.join(ROOT, "kspacejet_amd", "csrc"), "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "error_mapping_probe.cpp"), "-o", str(exe), "-L/opt/rocm/lib",
                    "-Wl,-rpath,/opt/rocm/lib", "-lamdhip64"], check=True)
    got, want = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")[:2]
    assert got.split() == want.split()
