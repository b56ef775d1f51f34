"""C-ABI checks that need no GPU: the built libraries load and export every
function the headers in include/ declare.  No compute call is made."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
PKG = os.path.join(ROOT, "kspacejet_amd")


def declared_functions(header: str, prefix: str):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"\w+)\s*\(", text)))


def exported(lib_path: str):
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_gpu_library_exports_every_declared_symbol():
    lib = os.path.join(PKG, "libksj_gpu.so")
    assert os.path.exists(lib), "run python -m kspacejet_amd.build"
    ctypes.CDLL(lib)  # loads without a GPU
    names = declared_functions("ksj_gpu.h", "ksj_gpu_")
    assert len(names) >= 20
    missing = [n for n in names if n not in exported(lib)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from kspacejet_amd import _lib
    names = declared_functions("ksj_gpu.h", "ksj_gpu_")
    assert sorted(_lib.SIGNATURES) == names


def test_library_loads_through_binding_without_gpu():
    from kspacejet_amd import _lib
    lib = _lib.load()
    # Pure host queries: workspace sizing needs no device.
    assert lib.ksj_gpu_ifft2_workspace_bytes(8, 256, 256) >= 2 * 8 * 256 * 256 * 8
    assert lib.ksj_gpu_nudft2_adjoint_workspace_bytes(2, 4, 4) >= 2 * 16 * 8


def test_kernels_are_built_for_gfx950():
    # The offload bundle of our own build names its target triple.
    data = open(os.path.join(PKG, "libksj_gpu.so"), "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data
