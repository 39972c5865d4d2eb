"""Build libmocca_amd.so (HIP for gfx950) in-tree with hipcc.

    python -m mocca_amd.build        (or: python mocca_amd/build.py)

Each source is compiled to its own object (in parallel, one hipcc process per
file) and the objects are linked into the shared library.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = [os.path.join(HERE, "csrc", f) for f in ("mocca_kernels.hip", "mocca_stencil.hip", "mocca_scf.hip",
                                              "mocca_capi.hip", "mocca_scf_capi.hip",
                                              "mocca_hfb.hip", "mocca_hfb_large.hip", "mocca_noalloc.hip",
                                              "mocca_v1.hip")]
OUT = os.path.join(HERE, "libmocca_amd.so")
OBJ = os.path.join(HERE, "_build")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # keep the compiler from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]


def build(verbose=False, jobs=None):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ, exist_ok=True)
    objs = [os.path.join(OBJ, os.path.splitext(os.path.basename(s))[0] + ".o") for s in SRC]

    def compile_one(pair):
        src, obj = pair
        cmd = [hipcc] + FLAGS + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        return subprocess.run(cmd, capture_output=True, text=True)

    jobs = jobs or max(1, min(len(SRC), os.cpu_count() or 1))
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(compile_one, zip(SRC, objs)))
    for src, r in zip(SRC, results):
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise subprocess.CalledProcessError(r.returncode, "hipcc -c " + src)
    tmp = OUT + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
