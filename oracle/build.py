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
SOURCES = ["bvh_ref.cpp", "trace_ref.cpp", "baseline.cpp", "baseline_render.cpp", "pathtrace_ref.cpp", "sdf_ref.cpp",
           "fetch_count.cpp"]
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


NATIVE_FLAGS = ["-O3", "-march=native", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                "-pthread"]


def build_native(out_dir, verbose=False):
    """The same library compiled for the host it runs on (-O3 -march=native), for timing the
    CPU baseline there. Written to out_dir; returns its path."""
    import concurrent.futures
    cxx = os.environ.get("CXX") or shutil.which("g++") or "g++"
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "liboracle_native.so")
    compile_flags = [f for f in NATIVE_FLAGS if f != "-shared"]

    def compile_one(src):
        obj = os.path.join(out_dir, os.path.splitext(src)[0] + ".o")
        cmd = [cxx] + compile_flags + ["-c", os.path.join(HERE, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("native oracle build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
        return obj

    with concurrent.futures.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    cmd = [cxx, "-shared", "-pthread"] + objs + ["-o", out + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("native oracle link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(out + ".tmp", out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
