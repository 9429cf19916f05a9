"""Build recipe for libpuckgpu.so (gfx950) -- compiles every HIP source in
puck_amd/csrc with hipcc into puck_amd/libpuckgpu.so."""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpuckgpu.so")
ARCH = os.environ.get("PUCK_GPU_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    # every a*b+c is two roundings unless fmaf() is written: the engine's
    # arithmetic must match the CPU restatement bit for bit
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def build_host_adapter(verbose=False):
    """C++ puck::Index adapters over the C ABI and their test drivers (g++)."""
    root = os.path.dirname(HERE)
    out = os.path.join(root, "tests", "cpp", "host_adapter_test")
    cmd = ["g++", "-O2", "-std=c++17", "-o", out, os.path.join(root, "tests", "cpp", "host_adapter_main.cpp"),
           os.path.join(HERE, "host", "gpu_puck_index.cpp"), "-L" + HERE, "-lpuckgpu",
           "-Wl,-rpath,$ORIGIN/../../puck_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    ins = os.path.join(root, "tests", "cpp", "insert_adapter_test")
    cmd = ["g++", "-O2", "-std=c++17", "-o", ins, os.path.join(root, "tests", "cpp", "insert_adapter_main.cpp"),
           os.path.join(HERE, "host", "gpu_puck_index.cpp"), "-L" + HERE, "-lpuckgpu",
           "-Wl,-rpath,$ORIGIN/../../puck_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return out


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    cmd = [hipcc] + HIPCC_FLAGS + ["-I", os.path.join(HERE, "..", "include"), "-o", tmp] + sources()
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    build_host_adapter(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
