"""Builds the in-tree native libraries with hipcc for gfx950.

    python -m kspacejet_amd.build            # rebuild everything
    python -m kspacejet_amd.build --changed  # only sources newer than their objects

Outputs (git-ignored, shipped with the tree):
    kspacejet_amd/libksj_gpu.so                    HIP kernels + the C ABI of include/ksj_gpu.h
    kspacejet_amd/libksj_provider_<module>.so      one KSpaceJet Provider module per
                                                   csrc/provider/mod_<module>.cpp (ksj_provider_query)
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
    # Matrix-core accumulators in ordinary VGPRs: the AGPR form makes the register
    # allocator copy accumulators between the two files on every gather batch.
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
    f"-I{os.path.join(ROOT, 'include')}",
]
CXX_FLAGS = ["-O2", "-std=c++20", "-fPIC", "-Wall", "-fvisibility=hidden", "-D__HIP_PLATFORM_AMD__", f"-I{os.path.join(ROOT, 'include')}",
             "-I/opt/rocm/include"]

GPU_LIB = os.path.join(PKG, "libksj_gpu.so")

def provider_lib(module: str) -> str:
    return os.path.join(PKG, f"libksj_provider_{module}.so")


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
    prov_headers = glob.glob(os.path.join(CSRC, "provider", "*.h*")) + glob.glob(os.path.join(CSRC, "provider", "*.inc"))
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
    # Every module links the shared provider core plus its own mod_<module>.cpp,
    # exports only ksj_provider_query, and finds libksj_gpu.so beside itself.
    core = [s for s in prov_srcs if not os.path.basename(s).startswith("mod_")]
    core_objs = [os.path.join(OBJ, "provider_" + os.path.basename(s) + ".o") for s in core]
    for src in prov_srcs:
        name = os.path.basename(src)
        if not name.startswith("mod_"):
            continue
        lib = provider_lib(name[len("mod_"):-len(".cpp")])
        objs = core_objs + [os.path.join(OBJ, "provider_" + name + ".o")]
        if not changed_only or _needs(lib, objs + [GPU_LIB]):
            _run(["g++", "-shared", "-fPIC", "-o", lib, *objs, f"-L{PKG}", "-lksj_gpu", "-Wl,-rpath,$ORIGIN",
                  "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-Wl,--no-undefined",
                  f"-Wl,--version-script={os.path.join(CSRC, 'provider', 'exports.map')}"])


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--changed", action="store_true")
    ap.add_argument("-j", "--jobs", type=int, default=8)
    args = ap.parse_args(argv)
    build(changed_only=args.changed, jobs=args.jobs)
    mods = sorted(glob.glob(os.path.join(PKG, "libksj_provider_*.so")))
    print(f"built {GPU_LIB}" + "".join(f"\n      {m}" for m in mods))
    return 0


if __name__ == "__main__":
    sys.exit(main())
