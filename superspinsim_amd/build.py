"""Compile the HIP library ``lib/libssmx.so`` for gfx950 (in-tree).

Each source is compiled to its own object in parallel (the kernels of one
file never call device code of another), then the objects are linked into
one shared library.
"""

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
OBJ_DIR = os.path.join(LIB_DIR, "_obj")
SOURCES = ["ssmx_propagate.hip", "ssmx_adjoint.hip", "ssmx_queue.hip",
           "ssmx_wide.hip", "ssmx_adjoint_wide.hip", "ssmx_dense.hip",
           "ssmx_chunks.hip"]

# -packed-fp32-ops: never emit v_pk_{add,mul,fma}_f32 (some of their forms
# have given wrong results on MI355X without any error); the host pass warns
# that it does not know the feature, which is harmless.
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc",
                 shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise FileNotFoundError("hipcc not found")


def _jobs():
    try:
        cpus = len(os.sched_getaffinity(0))
    except AttributeError:
        cpus = os.cpu_count() or 1
    return max(1, min(len(SOURCES), cpus, 16))


def build_library(verbose=False):
    """Compile every HIP source into ``lib/libssmx.so``; returns its path."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "libssmx.so")
    tmp = out + ".tmp"
    cc = hipcc()
    sources = list(SOURCES)

    def compile_one(src):
        obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
        cmd = [cc] + HIPCC_FLAGS + ["-c", "-o", obj,
                                    os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        return obj

    with ThreadPoolExecutor(max_workers=_jobs()) as pool:
        objs = list(pool.map(compile_one, sources))
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build_library(verbose=True))
