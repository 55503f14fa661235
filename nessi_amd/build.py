"""Build libcntrx.so (gfx950) in-tree: nessi_amd/lib/libcntrx.so."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib", "libcntrx.so")
SRC = os.path.join(HERE, "csrc", "cntrx_lib.hip")
DEPS = ["cntrx_lib.hip", "cntrx_device.h", "cntrx_jobs.h", "kernels_step.inc", "kernels_misc.inc",
        "kernels_rec1.inc", "kernels_rec1p.inc", "kernels_recmw.inc", "kernels_mm.inc", "kernels_zgemm.inc", "kernels_incr.inc", "lattice.inc",
        "server.inc",
        "textio.inc"]


def hipcc_command(out=LIB):
    return ["hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "csrc"),
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-o", out, SRC]


# C++ programs on the drop-in header include/cntr_amd.hpp (g++, linked to libcntrx.so)
CPP_TESTS = [("tests/cpp/dropin_driver.cpp", "tests/cpp/bin/dropin_driver")]


def cpp_command(src, out):
    return ["g++", "-O2", "-std=c++14", "-Wall", "-I" + os.path.join(ROOT, "include"), src,
            "-L" + os.path.dirname(LIB), "-lcntrx",
            "-Wl,-rpath,$ORIGIN/" + os.path.relpath(os.path.dirname(LIB), os.path.dirname(out)),
            "-o", out]


def build(verbose=False):
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = hipcc_command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    for src, out in CPP_TESTS:
        src, out = os.path.join(ROOT, src), os.path.join(ROOT, out)
        os.makedirs(os.path.dirname(out), exist_ok=True)
        cmd = cpp_command(src, out)
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB)
