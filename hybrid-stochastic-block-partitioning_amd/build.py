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
# uncontracted arithmetic); packed FP32 instructions off (the flag acts in the device pass; the
# host pass reports it as unknown and ignores it). check_no_packed_fp32 verifies the result.
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


LLVM = os.environ.get("SBPX_LLVM_BIN", "/opt/rocm/llvm/bin")
PACKED_FP32 = ("v_pk_add_f32", "v_pk_mul_f32", "v_pk_fma_f32")


def check_no_packed_fp32(lib=LIB):
    """Fails the build if the library's gfx950 code contains packed FP32 instructions: the device
    code objects (one bundle per source in the .hip_fatbin section) are unbundled and
    disassembled, and the listing searched."""
    import tempfile

    with tempfile.TemporaryDirectory() as tmp:
        fat = os.path.join(tmp, "fatbin.bin")
        subprocess.check_call([os.path.join(LLVM, "llvm-objcopy"), "--dump-section", f".hip_fatbin={fat}", lib])
        data = open(fat, "rb").read()
        magic = b"__CLANG_OFFLOAD_BUNDLE__"
        starts = []
        i = data.find(magic)
        while i != -1:
            starts.append(i)
            i = data.find(magic, i + 1)
        found = []
        for k, a in enumerate(starts):
            b = starts[k + 1] if k + 1 < len(starts) else len(data)
            bundle = os.path.join(tmp, f"b{k}.bin")
            code = os.path.join(tmp, f"b{k}.o")
            with open(bundle, "wb") as f:
                f.write(data[a:b])
            subprocess.check_call([os.path.join(LLVM, "clang-offload-bundler"), "--type=o",
                                   f"--targets=hipv4-amdgcn-amd-amdhsa--{ARCH}", f"--input={bundle}",
                                   f"--output={code}", "--unbundle"])
            if os.path.getsize(code) == 0:
                continue
            listing = subprocess.run([os.path.join(LLVM, "llvm-objdump"), "-d", code], capture_output=True,
                                     text=True, check=True).stdout
            found += [line.strip() for line in listing.splitlines() if any(op in line for op in PACKED_FP32)]
        if found:
            raise SystemExit("packed FP32 instructions in the device code:\n" + "\n".join(found[:20]))
        print(f"device code: {len(starts)} code-object bundles, no packed FP32 instructions", flush=True)


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
    check_no_packed_fp32()
    build_host()
    build_caller()


if __name__ == "__main__":
    build()
    sys.exit(0)
