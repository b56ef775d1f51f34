"""Builds the in-tree native libraries with hipcc for gfx950.

    python -m kspacejet_amd.build            # rebuild everything
    python -m kspacejet_amd.build --changed  # only sources newer than their objects

Outputs (git-ignored, shipped with the tree):
    kspacejet_amd/libksj_gpu.so           HIP kernels + the C ABI of include/ksj_gpu.h
    kspacejet_amd/libksj_provider_gpu.so  KSpaceJet Provider module (ksj_provider_query)
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
ARCH = os.environ.get("KSJ_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-Wall", "-Wno-unused-result",
    # No packed FP32 instructions anywhere in the device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{os.path.join(ROOT, 'include')}",
]
CXX_FLAGS = ["-O2", "-std=c++20", "-fPIC", "-Wall", "-D__HIP_PLATFORM_AMD__", f"-I{os.path.join(ROOT, 'include')}",
             "-I/opt/rocm/include"]

GPU_LIB = os.path.join(PKG, "libksj_gpu.so")
PROVIDER_LIB = os.path.join(PKG, "libksj_provider_gpu.so")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def _needs(obj, deps):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def build(changed_only: bool = False, jobs: int = 8) -> None:
    os.makedirs(OBJ, exist_ok=True)
    headers = glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(ROOT, "include", "*.h"))
    hip_srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    host_srcs = sorted(glob.glob(os.path.join(CSRC, "*.cpp")))
    prov_srcs = sorted(glob.glob(os.path.join(CSRC, "provider", "*.cpp")))
    prov_headers = glob.glob(os.path.join(CSRC, "provider", "*.h*"))
    tasks = []
    for src in hip_srcs:
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        if not changed_only or _needs(obj, [src] + headers):
            tasks.append([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
    for src in host_srcs:
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        if not changed_only or _needs(obj, [src] + headers):
            tasks.append([HIPCC, *HIP_FLAGS, "-x", "c++", "-c", src, "-o", obj])
    for src in prov_srcs:
        obj = os.path.join(OBJ, "provider_" + os.path.basename(src) + ".o")
        if not changed_only or _needs(obj, [src] + headers + prov_headers):
            tasks.append(["g++", *CXX_FLAGS, f"-I{CSRC}", "-c", src, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, tasks):
            if out.strip():
                print(out)
    gpu_objs = [os.path.join(OBJ, os.path.basename(s) + ".o") for s in hip_srcs + host_srcs]
    if not changed_only or _needs(GPU_LIB, gpu_objs):
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", GPU_LIB, *gpu_objs, "-lpthread"])
    prov_objs = [os.path.join(OBJ, "provider_" + os.path.basename(s) + ".o") for s in prov_srcs]
    if prov_objs and (not changed_only or _needs(PROVIDER_LIB, prov_objs + [GPU_LIB])):
        _run(["g++", "-shared", "-fPIC", "-o", PROVIDER_LIB, *prov_objs, f"-L{PKG}", "-lksj_gpu",
              "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib",
              "-Wl,--no-undefined"])


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--changed", action="store_true")
    ap.add_argument("-j", "--jobs", type=int, default=8)
    args = ap.parse_args(argv)
    build(changed_only=args.changed, jobs=args.jobs)
    print(f"built {GPU_LIB}" + (f" and {PROVIDER_LIB}" if os.path.exists(PROVIDER_LIB) else ""))
    return 0


if __name__ == "__main__":
    sys.exit(main())
