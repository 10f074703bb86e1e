"""Build step of the package: compiles the HIP library for gfx950 in-tree.

``python -m renderer_amd.build`` (or ``__graft_entry__.build()``) runs hipcc on
the sources under ``renderer_amd/csrc`` and writes ``renderer_amd/librt_hip.so``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librt_hip.so")

SOURCES = ["rt_kernels.hip", "rt_guide.hip", "rt_film.hip", "rt_mlt.hip", "rt_filter.hip", "rt_abi.cpp", "bvh_build.cpp", "scene_file.cpp",
           "film_io.cpp"]

# -packed-fp32-ops: keep the compiler from forming packed FP32 instructions.
# -amdgpu-mfma-vgpr-form (device code only): MFMA results in architectural
# VGPRs, so that the VALU reads them without copies out of the accumulation
# registers (the training kernel holds its gradient sums in registers and
# otherwise moved every other MFMA result through them).
HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Xarch_device", "-mllvm=-amdgpu-mfma-vgpr-form",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def build_library(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(HERE, "..", "include"), "-o", tmp] + srcs + ["-lpthread", "-lz"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building librt_hip.so")
    os.replace(tmp, LIB)
    return LIB


EXAMPLE = os.path.join(HERE, "host", "example_box")


def build_host_example(verbose=False):
    """The C++ host example (renderer_amd/host/example_box.cpp), linked
    against librt_hip.so with an rpath to the package directory."""
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(HERE, "..", "include"),
           os.path.join(HERE, "host", "example_box.cpp"), "-L", HERE, "-lrt_hip", "-Wl,-rpath,$ORIGIN/..",
           "-o", EXAMPLE]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building the host example")
    return EXAMPLE


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_host_example(verbose=True))
