"""Builds the CPU oracle (test infrastructure only) into oracle/liboracle.so with g++.

The oracle is compiled with -ffp-contract=off and without fast-math so that its binary32
arithmetic follows the written order, like the product kernels. -march is left at the baseline
x86-64 so the library built here runs on any host.
"""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
SOURCES = ["bvh_ref.cpp", "trace_ref.cpp", "baseline.cpp", "pathtrace_ref.cpp", "sdf_ref.cpp", "fetch_count.cpp"]
FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-pthread"]


def build(verbose=False):
    cxx = os.environ.get("CXX") or shutil.which("g++") or "g++"
    tmp = LIB + ".tmp"
    cmd = [cxx] + FLAGS + [os.path.join(HERE, s) for s in SOURCES] + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
