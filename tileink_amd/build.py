"""Builds libtileink_hip.so in-tree: the gfx950 code object from kernels.hip,
embedded into a host library compiled from the C++ runtime and recorder."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtileink_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

DEVICE_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-std=c++17", "-Wall", "-Wno-unused-function",
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-Wextra", "-Wno-unused-parameter",
              "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    hsaco = os.path.join(BUILD, "tileink_gfx950.hsaco")
    _run([os.path.join(ROCM, "bin", "hipcc"), "--genco", *DEVICE_FLAGS, "-I", INCLUDE, "-I", CSRC,
          os.path.join(CSRC, "kernels.hip"), "-o", hsaco])
    objs = []
    for src in ("canvas.cpp", "runtime.cpp"):
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        _run(["g++", *HOST_FLAGS, "-I", INCLUDE, "-I", CSRC, "-I", os.path.join(ROCM, "include"),
              "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    # The no-alloc scratch array: an ordinary HIP translation unit (host and device
    # parts), so that the HIP runtime registers and allocates its device array.
    scratch = os.path.join(BUILD, "scratch.o")
    _run([os.path.join(ROCM, "bin", "hipcc"), "-c", "-fPIC", "--no-offload-compress", *DEVICE_FLAGS, "-I", INCLUDE,
          "-I", CSRC, os.path.join(CSRC, "scratch.hip"), "-o", scratch])
    objs.append(scratch)
    blob = os.path.join(BUILD, "code_object.o")
    _run(["gcc", "-c", f'-DTK_HSACO_PATH="{hsaco}"', "-x", "assembler-with-cpp",
          os.path.join(CSRC, "code_object.S"), "-o", blob])
    objs.append(blob)
    tmp = LIB + ".tmp"
    _run(["g++", "-shared", "-o", tmp, *objs, "-L", os.path.join(ROCM, "lib"),
          f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}", "-lamdhip64"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
