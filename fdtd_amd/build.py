"""Build the engine library libfdtd_amd.so for gfx950 with hipcc.

Usage: python -m fdtd_amd.build  (or fdtd_amd.build.build()).
The library is built in-tree (fdtd_amd/libfdtd_amd.so) so that it travels with
the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfdtd_amd.so")
SOURCES = ["api.hip", "kernels.hip", "kernels_pair.hip", "fused.hip", "adjoint.hip", "scene.cpp", "port.cpp",
           "marching_loop.cpp"]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "--offload-arch=gfx950",
    # every a*b+c is a rounded multiply and a rounded add, in source order
    "-ffp-contract=off",
    # no packed FP32 instructions (v_pk_*_f32) in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def hipcc():
    path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")
    return path if os.path.exists(path) else "hipcc"


COPY_LIB = os.path.join(HERE, "libhbmcopy.so")


def _hipcc(out, srcs, verbose):
    tmp = out + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed building {os.path.basename(out)}")
    os.replace(tmp, out)


def build(verbose=False):
    """Build the engine library and the copy-kernel measurement library."""
    _hipcc(LIB, [os.path.join(CSRC, s) for s in SOURCES], verbose)
    _hipcc(COPY_LIB, [os.path.join(CSRC, "hbm_copy.hip")], verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
