"""Build recipe for the test oracle (test infrastructure only).

  oracle/liborcl.so            g++ build of gpma_oracle.cpp (the CPU
                               restatement of the reference's update and BFS path)
  oracle/_ref/ref_cpu_baseline  g++ build of ref_cpu_baseline.cpp against the
                               reference's own cpu_baseline headers, compiled
                               where they lie under /root/reference (only when
                               that tree is present; the built binary travels
                               with the repository snapshot, the sources do not)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF_ROOT = "/root/reference"
REF_CPU_BASELINE = os.path.join(REF_ROOT, "cpu_baseline")
ORACLE_LIB = os.path.join(HERE, "liborcl.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_BIN = os.path.join(REF_DIR, "ref_cpu_baseline")


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("oracle build failed: " + " ".join(cmd) + "\n" + p.stdout)


def build():
    # OpenMP: the restatement runs the reference CPU path's parallel loops
    # (locate, windows, batch sort), so it can also be timed on all host cores
    _run(["g++", "-O2", "-std=c++17", "-fopenmp", "-shared", "-fPIC", os.path.join(HERE, "gpma_oracle.cpp"), "-o",
          ORACLE_LIB])
    built_ref = False
    if os.path.isdir(REF_CPU_BASELINE):
        os.makedirs(REF_DIR, exist_ok=True)
        # cpu_baseline/Makefile:2 builds with g++ -O3 -std=c++11
        _run(["g++", "-O3", "-std=c++11", f"-I{REF_CPU_BASELINE}", f"-I{REF_CPU_BASELINE}/algorithms",
              f"-I{REF_CPU_BASELINE}/containers", os.path.join(HERE, "ref_cpu_baseline.cpp"), "-o", REF_BIN])
        built_ref = True
    return ORACLE_LIB, (REF_BIN if built_ref else None)


if __name__ == "__main__":
    print(build())
    sys.exit(0)
