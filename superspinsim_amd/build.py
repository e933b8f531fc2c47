"""Compile the HIP library ``lib/libssmx.so`` for gfx950 (in-tree)."""

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
SOURCES = ["ssmx_propagate.hip", "ssmx_adjoint.hip", "ssmx_queue.hip",
           "ssmx_wide.hip",
           "ssmx_adjoint_wide.hip"]

# -packed-fp32-ops: never emit v_pk_{add,mul,fma}_f32 (see DESIGN notes on
# packed FP32 on this hardware); the host pass warns that it does not know
# the feature, which is harmless.
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
               "-shared", "-Xclang", "-target-feature", "-Xclang",
               "-packed-fp32-ops"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc",
                 shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise FileNotFoundError("hipcc not found")


def build_library(verbose=False):
    """Compile every HIP source into ``lib/libssmx.so``; returns its path."""
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "libssmx.so")
    tmp = out + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp] + \
        [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build_library(verbose=True))
