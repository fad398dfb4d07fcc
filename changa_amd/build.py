"""Build the native libraries in-tree.

  libchanga_host.so  g++ (C++17): tree build, node geometry, interaction lists,
                     Ewald h-loop table -- the caller side of the GPU boundary.
  libchanga_gpu.so   hipcc for gfx950: the HIP kernels and the extern "C"
                     HostCUDA/EwaldCUDA boundary (include/changa_gpu.h).

Run as `python -m changa_amd.build` or through __graft_entry__.build().
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")

HOST_LIB = os.path.join(PKG, "libchanga_host.so")
GPU_LIB = os.path.join(PKG, "libchanga_gpu.so")

HOST_SRC = ["host/host_tree.cpp", "host/host_walk.cpp", "host/host_api.cpp"]
GPU_SRC = ["gpu/gravity.hip", "gpu/batch.hip", "gpu/adjoint.hip", "gpu/ewald.hip", "gpu/moments.hip",
           "gpu/walk.hip", "gpu/tree.hip", "gpu/fused.hip", "gpu/hostcuda.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instruction selection stays off for every HIP object.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_host():
    srcs = [os.path.join(CSRC, s) for s in HOST_SRC]
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-Wall",
          "-I" + os.path.join(CSRC, "host"), "-o", HOST_LIB] + srcs)


def build_gpu():
    srcs = [os.path.join(CSRC, s) for s in GPU_SRC]
    _run([HIPCC] + HIP_FLAGS + ["-I" + INC, "-I" + os.path.join(CSRC, "gpu"),
                                "-o", GPU_LIB] + srcs)


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all(gpu=True):
    build_host()
    if gpu:
        build_gpu()
    build_oracle()


if __name__ == "__main__":
    build_all(gpu="--no-gpu" not in sys.argv)
