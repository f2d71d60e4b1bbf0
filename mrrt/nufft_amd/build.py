"""Build ``libmi355nufft.so`` (HIP, gfx950) in-tree with hipcc.

Usage: ``python -m mrrt.nufft_amd.build`` or ``build_library()``.
The library is written next to this file and loaded by ``_lib.py``.
"""

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_NAME = "libmi355nufft.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

# translation units: (source, extra defines).  The tiled-kernel dispatch is
# compiled once per value type and kernel kind so the heaviest template
# instantiations build in parallel.
UNITS = [
    ("mi355nufft.hip", []),
    ("fft.hip", []),
    ("tiled.hip", ["-DMI355_TILED_T=float", "-DMI355_TILED_CK=0"]),
    ("tiled.hip", ["-DMI355_TILED_T=float", "-DMI355_TILED_CK=1"]),
    ("tiled.hip", ["-DMI355_TILED_T=double", "-DMI355_TILED_CK=0"]),
    ("tiled.hip", ["-DMI355_TILED_T=double", "-DMI355_TILED_CK=1"]),
]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed-FP32 (v_pk_*_f32) generation off
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc_path():
    p = os.path.join(ROCM, "bin", "hipcc")
    return p if os.path.exists(p) else shutil.which("hipcc")


def _filter(out):
    return "\n".join(l for l in out.splitlines() if "packed-fp32-ops' is not a recognized feature" not in l)


def _run(cmd):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return res.returncode, _filter(res.stdout)


def build_library(verbose=True, jobs=None):
    """Compile every translation unit (in parallel) and link the library."""
    hipcc = hipcc_path()
    if hipcc is None:
        raise RuntimeError("hipcc not found; cannot build libmi355nufft.so")
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    inc = "-I" + os.path.join(ROCM, "include")
    cmds, objs = [], []
    for src, defs in UNITS:
        tag = "".join(d.split("=")[-1] for d in defs)
        obj = os.path.join(bdir, os.path.splitext(src)[0] + ("_" + tag if tag else "") + ".o")
        objs.append(obj)
        cmds.append([hipcc] + HIPCC_FLAGS + defs + [inc, "-c", os.path.join(CSRC, src), "-o", obj])
    jobs = jobs or min(len(cmds), max(1, os.cpu_count() or 1))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_run, cmds))
    for cmd, (rc, out) in zip(cmds, results):
        if verbose:
            print(" ".join(cmd), flush=True)
            if out.strip():
                print(out)
        if rc != 0:
            raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + out)
    tmp = LIB_PATH + ".tmp%d" % os.getpid()
    link = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + [
        "-o", tmp, "-L" + os.path.join(ROCM, "lib"), "-lrocfft", "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    if verbose:
        print(" ".join(link), flush=True)
    rc, out = _run(link)
    if rc != 0:
        if os.path.exists(tmp):
            os.remove(tmp)
        raise RuntimeError("link failed:\n" + out)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    build_library(verbose=True)
    sys.exit(0)
