"""Build recipe for the gfx950 library (libgpma.so, in this directory).

hipcc compiles each .hip translation unit for gfx950 only, with packed FP32
instructions disabled; gen.cpp and serve.cpp are plain host C++ (g++). Objects go to
_build/ beside this file and the shared library lands next to the package.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgpma.so")
DEMO = os.path.join(HERE, "demo", "gpma_bfs_demo")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.environ.get("HIPCC", os.path.join(ROCM, "bin", "hipcc"))
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
HIP_SOURCES = ["sort.hip", "msd.hip", "update.hip", "bfs.hip", "msbfs.hip", "gpma.hip"]
CXX_SOURCES = ["gen.cpp", "serve.cpp"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + p.stdout)
    return p.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        objs.append(obj)
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    for src in CXX_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        objs.append(obj)
        # host C++; serve.cpp calls the HIP runtime API (streams, events, copies)
        jobs.append(["g++", "-O3", "-std=c++17", "-fPIC", "-pthread", "-Wall", f"-I{ROCM}/include",
                     "-D__HIP_PLATFORM_AMD__", "-c", os.path.join(CSRC, src), "-o", obj])
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp, "-lpthread"])
    os.replace(tmp, LIB)
    # the reference demo's protocol against the C++ facade (include/gpma.hpp)
    _run(["g++", "-O2", "-std=c++17", f"-I{os.path.join(os.path.dirname(HERE), 'include')}",
          os.path.join(HERE, "demo", "gpma_bfs_demo.cpp"), "-o", DEMO, f"-L{HERE}", "-lgpma",
          f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN"])
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
