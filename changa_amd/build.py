"""Build the native libraries in-tree.

  libchanga_host.so  g++ (C++17): tree build, node geometry, interaction lists,
                     Ewald h-loop table -- the caller side of the GPU boundary.
  libchanga_gpu.so   hipcc for gfx950: the HIP kernels and the extern "C"
                     HostCUDA/EwaldCUDA boundary (include/changa_gpu.h).
  libchanga_gpu_ltw.so  the same HostCUDA/EwaldCUDA boundary built on the
                     wire layouts of ChaNGa's GPU_LOCAL_TREE_WALK build, whose
                     TreePieceCellListDataTransferLocal runs the local tree
                     walk (gpu/ltw.hip).

Run as `python -m changa_amd.build` or through __graft_entry__.build().
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")

HOST_LIB = os.path.join(PKG, "libchanga_host.so")
GPU_LIB = os.path.join(PKG, "libchanga_gpu.so")
LTW_LIB = os.path.join(PKG, "libchanga_gpu_ltw.so")

HOST_SRC = ["host/host_tree.cpp", "host/host_walk.cpp", "host/host_api.cpp"]
GPU_SRC = ["gpu/gravity.hip", "gpu/batch.hip", "gpu/adjoint.hip", "gpu/ewald.hip", "gpu/moments.hip",
           "gpu/walk.hip", "gpu/tree.hip", "gpu/fused.hip", "gpu/hostcuda.hip"]
# the boundary's sources, compiled again with -DGPU_LOCAL_TREE_WALK
LTW_SRC = ["gpu/gravity.hip", "gpu/ewald.hip", "gpu/hostcuda.hip", "gpu/ltw.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instruction selection stays off for every HIP object.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function"]
OBJ_DIR = os.path.join(PKG, "_build")
JOBS = max(1, min(len(GPU_SRC), os.cpu_count() or 1, 16))


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_host():
    srcs = [os.path.join(CSRC, s) for s in HOST_SRC]
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-Wall",
          "-I" + os.path.join(CSRC, "host"), "-o", HOST_LIB] + srcs)


def build_gpu():
    """One object per source (both libraries' objects compiled in parallel),
    then one shared link per library."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    inc = ["-I" + INC, "-I" + os.path.join(CSRC, "gpu")]
    main, ltw = [], []
    for srcs, prefix, extra, objs in ((GPU_SRC, "", [], main),
                                      (LTW_SRC, "ltw_", ["-DGPU_LOCAL_TREE_WALK"], ltw)):
        for src in srcs:
            obj = os.path.join(OBJ_DIR, prefix + os.path.basename(src).replace(".hip", ".o"))
            objs.append((obj, [HIPCC] + HIP_FLAGS + extra + inc +
                         ["-c", os.path.join(CSRC, src), "-o", obj]))
    with ThreadPoolExecutor(JOBS) as ex:
        list(ex.map(lambda j: _run(j[1]), main + ltw))
    for lib, objs in ((GPU_LIB, main), (LTW_LIB, ltw)):
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", lib] + [o for o, _ in objs])


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all(gpu=True):
    build_host()
    if gpu:
        build_gpu()
    build_oracle()


if __name__ == "__main__":
    build_all(gpu="--no-gpu" not in sys.argv)
