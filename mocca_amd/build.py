"""Build libmocca_amd.so (HIP for gfx950) in-tree with hipcc.

    python -m mocca_amd.build        (or: python mocca_amd/build.py)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = [os.path.join(HERE, "csrc", f) for f in ("mocca_kernels.hip", "mocca_stencil.hip", "mocca_scf.hip",
                                              "mocca_capi.hip", "mocca_scf_capi.hip",
                                              "mocca_hfb.hip", "mocca_hfb_large.hip", "mocca_noalloc.hip")]
OUT = os.path.join(HERE, "libmocca_amd.so")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # keep the compiler from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = OUT + ".tmp"
    cmd = [hipcc] + FLAGS + SRC + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
