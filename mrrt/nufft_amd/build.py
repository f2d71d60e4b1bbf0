"""Build ``libmi355nufft.so`` (HIP, gfx950) in-tree with hipcc.

Usage: ``python -m mrrt.nufft_amd.build`` or ``build_library()``.
The library is written next to this file and loaded by ``_lib.py``.
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_NAME = "libmi355nufft.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

SOURCES = ["mi355nufft.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed-FP32 (v_pk_*_f32) generation off
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc_path():
    p = os.path.join(ROCM, "bin", "hipcc")
    return p if os.path.exists(p) else shutil.which("hipcc")


def build_library(verbose=True):
    hipcc = hipcc_path()
    if hipcc is None:
        raise RuntimeError("hipcc not found; cannot build libmi355nufft.so")
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB_PATH + ".tmp%d" % os.getpid()
    cmd = [hipcc] + HIPCC_FLAGS + ["-I" + os.path.join(ROCM, "include")] + srcs + [
        "-o", tmp, "-L" + os.path.join(ROCM, "lib"), "-lrocfft",
        "-Wl,-rpath," + os.path.join(ROCM, "lib"),
    ]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if res.returncode != 0:
        if os.path.exists(tmp):
            os.remove(tmp)
        raise RuntimeError("hipcc failed:\n" + out)
    if verbose and out.strip():
        print(out)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    build_library(verbose=True)
    sys.exit(0)
