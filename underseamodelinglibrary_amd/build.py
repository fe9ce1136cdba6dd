"""Build libwq3d.so (HIP, gfx950) in-tree.

Flags: -ffp-contract=off keeps every FP64 expression in the order written
(no fused multiply-add contraction), which the bit-level parity with the
reference arithmetic relies on; packed FP32 instructions are disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "wq3d.hip")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libwq3d.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


EMU_LIB = os.path.join(OUT_DIR, "libwq3d_emu.so")


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_emulation(asan=True, verbose=False):
    """CPU build of the same kernel bodies (serial loops, host memory) for
    tests that check every access with AddressSanitizer.  Never loaded by
    the package itself."""
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = EMU_LIB + ".tmp"
    flags = ["--offload-host-only", "-O1", "-g", "-std=c++17", "-fPIC", "-shared",
             "-ffp-contract=off", "-fno-fast-math", "-DWQ3D_EMULATE"]
    if subprocess.run(["grep", "-qw", "fma", "/proc/cpuinfo"]).returncode == 0:
        flags.append("-mfma")
    if asan:
        flags += ["-Xarch_host", "-fsanitize=address", "-fno-omit-frame-pointer"]
    cmd = [HIPCC, *flags, "-I", os.path.join(ROOT, "include"), "-o", tmp, SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, EMU_LIB)
    return EMU_LIB


CPP_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "wave_queue_eigenray_basic.cpp")
CPP_TEST = os.path.join(OUT_DIR, "wave_queue_eigenray_basic")


def build_cpp_test(lib=LIB, out=CPP_TEST, extra=(), verbose=False):
    """The C++ wave_queue mirror's test program, linked against `lib`
    (libwq3d.so, or the emulation build for CPU runs, which also needs its
    sanitizer runtime in `extra`)."""
    os.makedirs(os.path.dirname(out), exist_ok=True)
    lib_dir = os.path.dirname(os.path.abspath(lib))
    name = os.path.basename(lib)[3:-3]
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-o", out + ".tmp", CPP_TEST_SRC,
           "-L" + lib_dir, "-l" + name, "-Wl,-rpath," + lib_dir, *extra]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
    sys.exit(0)
