"""Build the engine library libfdtd_amd.so for gfx950 with hipcc.

Usage: python -m fdtd_amd.build  (or fdtd_amd.build.build()).
Each source compiles to an object of its own (in parallel, under
fdtd_amd/_build/), then the objects link into fdtd_amd/libfdtd_amd.so, in-tree,
so that the library travels with the source tree.  Every call rebuilds.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfdtd_amd.so")
SOURCES = ["api.hip", "kernels.hip", "kernels_pair.hip", "adjoint.hip", "scene.cpp", "port.cpp",
           "marching_loop.cpp"]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
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


def _run(cmd, what, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed building {what}")


def _link(out, srcs, verbose):
    """Compile srcs to objects in parallel, then link them into out."""
    os.makedirs(OBJ, exist_ok=True)
    objs = [os.path.join(OBJ, os.path.basename(s) + ".o") for s in srcs]
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as pool:
        list(pool.map(lambda so: _run([hipcc()] + HIPCC_FLAGS + ["-c", "-o", so[1], so[0]],
                                      os.path.basename(so[0]), verbose), zip(srcs, objs)))
    tmp = out + ".tmp"
    _run([hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs, os.path.basename(out), verbose)
    os.replace(tmp, out)


def build(verbose=False):
    """Build the engine library and the copy-kernel measurement library."""
    _link(LIB, [os.path.join(CSRC, s) for s in SOURCES], verbose)
    _link(COPY_LIB, [os.path.join(CSRC, "hbm_copy.hip")], verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
