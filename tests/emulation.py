"""CPU emulation build of libwq3d (test infrastructure): the same kernel
bodies run in serial loops over host memory, built with AddressSanitizer so
every array access of the kernels is bounds checked on the host."""
import ctypes as C
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def asan_runtime():
    out = subprocess.run(["/opt/rocm/bin/amdclang++", "-print-file-name=libclang_rt.asan-x86_64.so"],
                         capture_output=True, text=True, check=True).stdout.strip()
    return out


def load():
    from underseamodelinglibrary_amd import _lib, build
    if not os.path.exists(build.EMU_LIB):
        build.build_emulation()
    return _lib.bind(C.CDLL(build.EMU_LIB))


def run_script_under_asan(script, timeout=600):
    """Run a python snippet in a child process with the ASan runtime preloaded."""
    env = dict(os.environ)
    env["LD_PRELOAD"] = asan_runtime()
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=0"
    env["PYTHONPATH"] = os.pathsep.join([ROOT, os.path.join(ROOT, "oracle"),
                                         os.path.join(ROOT, "tests")])
    return subprocess.run([sys.executable, "-c", script], env=env, capture_output=True,
                          text=True, timeout=timeout)
