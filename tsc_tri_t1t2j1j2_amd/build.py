"""Build libtsc_hotpath.so (gfx950) in-tree with hipcc.

The shared library is the C-ABI declared in include/tsc_hotpath.h. It is
built next to this file so that it travels with the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtsc_hotpath.so")
SOURCES = ["tsc_capi.hip", "tsc_grouped.hip", "tsc_pk.hip", "tsc_gd.hip", "tsc_vec.hip", "tsc_rho.hip", "tsc_planner.hip", "tsc_loop.hip"]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIPFLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # keep packed FP32 instructions out of the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unused-result", "-Wno-unused-value",
]


def build(verbose=False):
    """Compile every HIP source and link the shared library."""
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    objs = []
    os.makedirs(os.path.join(HERE, "_obj"), exist_ok=True)
    cmds = []
    for src in SOURCES:
        obj = os.path.join(HERE, "_obj", src.replace(".hip", ".o"))
        cmd = [hipcc, *HIPFLAGS, "-I", os.path.join(ROOT, "include"), "-c",
               os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        cmds.append(cmd)
        objs.append(obj)
    # the sources compile independently: a few at a time
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=min(4, len(cmds))) as pool:
        for r in list(pool.map(lambda c: subprocess.run(c, check=False), cmds)):
            if r.returncode != 0:
                raise subprocess.CalledProcessError(r.returncode, r.args)
    tmp = LIB + ".tmp"
    cmd = [hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp,
           "-L", os.path.join(ROCM, "lib"), "-lrocsolver", "-lrocblas",
           "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
