"""Write golden table1 vectors with the reference's own C kernels.

Inputs (tables, tm, grids, samples) are seeded; outputs come from
``oracle/_ref/libnufft_table_ref.so`` -- mrrt/nufft/_extensions/c/nufft_table.c
compiled from the reference tree by ``oracle/build_ref.sh`` -- driven the way
the Cython bridge drives it (_extensions/_nufft_table.pyx:29-704).

Run in the build container (where /root/reference exists):
    sh oracle/build_ref.sh && python tests/golden/make_golden.py
Output: tests/golden/table1_golden.npz (inputs and expected outputs only).
"""

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import nufft_oracle as O  # noqa: E402
from oracle import ref_ctypes as R  # noqa: E402

CASES = [
    # (ndim, Nd, Kd, Jd, L)
    (1, (20,), (40,), (6,), 256),
    (1, (21,), (33,), (7,), 512),
    (2, (16, 16), (32, 33), (6, 7), 256),
    (2, (12, 10), (24, 20), (4, 5), 1024),
    (3, (6, 5, 7), (12, 10, 14), (6, 5, 4), 128),
]
M = 96
NREPS = 2


def main():
    if not R.available():
        raise SystemExit("oracle/_ref/libnufft_table_ref.so is missing: run oracle/build_ref.sh")
    rng = np.random.default_rng(20240601)
    out = {}
    for ci, (ndim, Nd, Kd, Jd, L) in enumerate(CASES):
        for prec, rdt in (("single", np.float32), ("double", np.float64)):
            cdt = np.result_type(rdt, np.complex64)
            for ph in ("real", "complex"):
                key = "c%d_%s_%s" % (ci, prec, ph)
                hs = []
                for d in range(ndim):
                    h, _ = O.table_make1("fast", Nd[d], Jd[d], Kd[d], L, ph)
                    hs.append(h.astype(cdt) if ph == "complex" else np.real(h).astype(rdt))
                # points: most in [-K/2, K/2), some beyond one period either way
                tm = np.empty((M, ndim), rdt)
                for d in range(ndim):
                    t = rng.uniform(-Kd[d] / 2, Kd[d] / 2, M)
                    t[:8] = rng.uniform(-1.5 * Kd[d], 1.5 * Kd[d], 8)
                    tm[:, d] = t.astype(rdt)
                nk = int(np.prod(Kd))
                ck = (rng.standard_normal((nk, NREPS)) + 1j * rng.standard_normal((nk, NREPS))).astype(cdt)
                fm_in = (rng.standard_normal((M, NREPS)) + 1j * rng.standard_normal((M, NREPS))).astype(cdt)
                out[key + "_fwd"] = R.ref_interp_forward(ck, Kd, hs, Jd, L, tm)
                out[key + "_adj"] = R.ref_interp_adjoint(fm_in, Kd, hs, Jd, L, tm)
                out[key + "_tm"] = tm
                out[key + "_ck"] = ck
                out[key + "_fm"] = fm_in
                for d in range(ndim):
                    out[key + "_h%d" % d] = hs[d]
                out[key + "_geom"] = np.array([ndim, L] + list(Kd) + list(Jd), dtype=np.int64)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "table1_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes,", len(out), "arrays")


if __name__ == "__main__":
    main()
