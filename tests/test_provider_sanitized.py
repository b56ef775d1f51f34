"""The Provider core under the host sanitizers (CPU only).

Every Provider module (the shared core that parses canonical JSON configs and walks
caller-supplied lease views, plus each module's descriptor table) is built again with
g++ -fsanitize=address,undefined into a temporary directory, and the CPU Provider tests
(ABI layout, config grammar, lease rejections of every module) and the config mutation
fuzz run against that build in a subprocess with the ASan runtime preloaded.  Any
AddressSanitizer or UndefinedBehaviorSanitizer report fails the test.  The GPU library
the modules link is the ordinary build; firings that would reach a device are not run."""
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "kspacejet_amd")
PROVIDER_SRC = os.path.join(PKG, "csrc", "provider")
TESTS = ["tests/test_provider_abi.py", "tests/test_provider_conditioning.py", "tests/test_provider_calibration.py",
         "tests/test_provider_fuzz.py"]
SANITIZE = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]


def _run(cmd, **kw):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, **kw)
    return r.returncode, r.stdout


def test_provider_modules_are_clean_under_asan_and_ubsan(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    if not os.path.exists(os.path.join(PKG, "libksj_gpu.so")):
        pytest.skip("libksj_gpu.so is not built")
    rc, asan = _run(["g++", "-print-file-name=libasan.so"])
    asan = asan.strip()
    if rc != 0 or not os.path.isabs(asan) or not os.path.exists(asan):
        pytest.skip("no ASan runtime for g++")
    flags = ["-O1", "-g", "-std=c++20", "-fPIC", "-fvisibility=hidden", "-D__HIP_PLATFORM_AMD__",
             f"-I{os.path.join(ROOT, 'include')}", "-I/opt/rocm/include", f"-I{os.path.join(PKG, 'csrc')}", *SANITIZE]
    srcs = sorted(glob.glob(os.path.join(PROVIDER_SRC, "*.cpp")))
    objs = {s: str(tmp_path / (os.path.basename(s) + ".o")) for s in srcs}
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        for src, (code, out) in zip(srcs, ex.map(lambda s: _run(["g++", *flags, "-c", s, "-o", objs[s]]), srcs)):
            assert code == 0, f"{src}:\n{out}"
    core = [objs[s] for s in srcs if not os.path.basename(s).startswith("mod_")]
    for src in srcs:
        name = os.path.basename(src)
        if not name.startswith("mod_"):
            continue
        lib = str(tmp_path / f"libksj_provider_{name[len('mod_'):-len('.cpp')]}.so")
        code, out = _run(["g++", "-shared", *SANITIZE, "-o", lib, *core, objs[src], f"-L{PKG}", "-lksj_gpu",
                          f"-Wl,-rpath,{PKG}", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib",
                          f"-Wl,--version-script={os.path.join(PROVIDER_SRC, 'exports.map')}"])
        assert code == 0, out
    env = dict(os.environ)
    env.update(LD_PRELOAD=asan, KSJ_PROVIDER_DIR=str(tmp_path),
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:detect_odr_violation=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    code, out = _run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                      "-k", "not without_a_device", *TESTS], cwd=ROOT, env=env, timeout=900)
    tail = "\n".join(out.splitlines()[-40:])
    assert "AddressSanitizer" not in out and "runtime error:" not in out, tail
    assert code == 0, tail
    assert " passed" in out, tail
