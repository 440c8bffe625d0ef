"""Build libgdpd.so (hand-written gfx950 HIP) in-tree with hipcc.

The library is the C-ABI drop-in (include/gdpd.h).  It is compiled for gfx950
only; packed-FP32 instructions are disabled (-packed-fp32-ops) so hipcc emits
plain FP32 arithmetic.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgdpd.so")
SOURCES = ["gdpd_kernels.hip", "gdpd_api.hip", "gdpd_plugins.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"


def hipcc_command(out=LIB):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    return [
        HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"),
        "-o", out,
    ] + srcs


CPP_DIR = os.path.join(ROOT, "tests", "cpp")
# C++ programs over the plugin shells used by the tests (link libgdpd.so)
DRIVERS = ["plugin_driver", "data_roundtrip", "deck_runner"]
DRIVER_SRC = os.path.join(CPP_DIR, "plugin_driver.cpp")


def build_driver(verbose=False):
    for name in DRIVERS:
        src = os.path.join(CPP_DIR, name + ".cpp")
        if not os.path.exists(src):
            continue
        cmd = ["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), src, "-o",
               os.path.join(CPP_DIR, name), "-L", HERE, "-lgdpd", f"-Wl,-rpath,{HERE}",
               "-Wl,-rpath,$ORIGIN/../../gpumdpd_amd"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"g++ failed building {name}")


def build(verbose=False):
    cmd = hipcc_command(LIB + ".tmp")
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libgdpd.so")
    os.replace(LIB + ".tmp", LIB)
    if os.path.exists(DRIVER_SRC):
        build_driver(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
