"""Builds libfamehip.so in-tree for gfx950 (hipcc cross-compiles without a GPU).

    python -m fame_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libfamehip.so")
INDEX_LIB = os.path.join(HERE, "libfameindex.so")
SOURCES = [os.path.join(HERE, "csrc", "fame_device.hip"), os.path.join(HERE, "csrc", "fame_paired.hip"),
           os.path.join(HERE, "csrc", "fame_host.cpp"), os.path.join(HERE, "csrc", "fame_serve.cpp"),
           os.path.join(HERE, "csrc", "fame_fastq.cpp")]
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17", "-pthread",
         # packed FP32 instructions are kept out of all device code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    cmd = [hipcc] + FLAGS + ["-o", tmp] + SOURCES + ["-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libfamehip.so")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, LIB)
    # host-only index builder (C++ / OpenMP)
    tmp = INDEX_LIB + ".tmp"
    r = subprocess.run(["g++", "-O3", "-fPIC", "-shared", "-std=c++17", "-fopenmp", "-o", tmp,
                        os.path.join(HERE, "csrc", "fame_index.cpp")], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building libfameindex.so")
    os.replace(tmp, INDEX_LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
