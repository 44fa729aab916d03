"""Build libcnpc_hip.so (gfx950) in-tree with hipcc.

Usage: python -m cosmonpc_amd.build_lib   (or cosmonpc_amd.build_lib.build())
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["cnpc_api.hip", "paint.hip", "spectral.hip", "contract.hip", "synth.hip", "survey.hip", "zgemm.hip", "fft400.hip"]
LIB = os.path.join(HERE, "libcnpc_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def hipcc_command(out=LIB):
    return [
        os.path.join(ROCM, "bin", "hipcc"),
        "--offload-arch=gfx950",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-munsafe-fp-atomics",
        # packed FP32 instructions are kept off for all device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(HERE, "..", "include"),
        "-o", out,
        *[os.path.join(CSRC, s) for s in SOURCES],
        "-L", os.path.join(ROCM, "lib"),
        "-lrocfft",
        "-Wl,-rpath," + os.path.join(ROCM, "lib"),
    ]


def build(verbose=True):
    cmd = hipcc_command(LIB + ".tmp")
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
