"""Builds libsbpx.so (HIP kernels + C ABI) for gfx950 and libsbp_host.so (the C++ layer that
restores the reference's API on top of the C ABI), in-tree, with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
LIB = os.path.join(HERE, "libsbpx.so")
HOST_LIB = os.path.join(HERE, "libsbp_host.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SBPX_ARCH", "gfx950")

DEVICE_SOURCES = ["sort_scan.hip", "blockmodel.hip", "sweep.hip", "mh.hip", "merge.hip", "mlp.hip", "learned.hip", "cells.hip", "api.cpp",
                  "generator.cpp"]
HOST_SOURCES = ["sbp_host.cpp", "graph_io.cpp", "batch.cpp", "capi.cpp"]

# No FMA contraction (the objective's FP32 cell chain and FP64 sums must match the reference's
# uncontracted arithmetic); packed FP32 instructions off.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_device():
    srcs = [os.path.join(CSRC, s) for s in DEVICE_SOURCES]
    tmp = LIB + ".tmp"
    run([HIPCC] + HIP_FLAGS + ["-shared", "-x", "hip"] + srcs + ["-o", tmp])
    os.replace(tmp, LIB)


def build_host():
    srcs = [os.path.join(HOST, s) for s in HOST_SOURCES if os.path.exists(os.path.join(HOST, s))]
    if not srcs:
        return
    tmp = HOST_LIB + ".tmp"
    run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off", f"-I{os.path.join(ROOT, 'include')}",
         f"-I{HOST}"] + srcs + ["-o", tmp, f"-L{HERE}", "-lsbpx", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOST_LIB)


CALLER_SRC = os.path.join(ROOT, "tests", "cpp", "toy_caller.cpp")
CALLER_BIN = os.path.join(ROOT, "tests", "cpp", "toy_caller")


def build_caller():
    """A reference-style C++ caller of the host layer, for the tests."""
    if not os.path.exists(CALLER_SRC):
        return
    tmp = CALLER_BIN + ".tmp"
    run(["g++", "-O2", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}", f"-I{HOST}", CALLER_SRC, "-o", tmp,
         f"-L{HERE}", "-lsbp_host", "-lsbpx", "-Wl,-rpath,$ORIGIN/../../hybrid-stochastic-block-partitioning_amd"])
    os.replace(tmp, CALLER_BIN)


def build():
    build_device()
    build_host()
    build_caller()


if __name__ == "__main__":
    build()
    sys.exit(0)
