"""Shared test inputs: the reference's own test trajectories and helpers."""

import numpy as np

GOLDEN_PATH = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "table1_golden.npz")


def uniform_freqs(Nd):
    """mrrt/nufft/tests/test_dtft.py:25-36."""
    if np.isscalar(Nd):
        Nd = (Nd,)
    fs = [2 * np.pi * np.arange(Nd[d]) / Nd[d] for d in range(len(Nd))]
    fs = np.meshgrid(*fs, indexing="ij")
    return np.hstack([f.reshape((-1, 1), order="F") for f in fs])


def perturbed_gridpoints(Nd, rel_std=0.5, seed=1234):
    """mrrt/nufft/tests/test_nufft.py:20-44."""
    rstate = np.random.RandomState(seed)
    if np.isscalar(Nd):
        Nd = (Nd,)
    Nd = np.asarray(Nd)
    omega = uniform_freqs(Nd)
    df = 2 * np.pi / Nd
    npoints = omega.shape[0]
    for d in range(len(Nd)):
        omega[:, d] += df[d] * rel_std * rstate.rand(npoints)
        omega[:, d] += np.min(omega[:, d])
        omega[:, d] *= (2 * np.pi) / np.max(omega[:, d])
    return omega


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def golden_cases():
    g = np.load(GOLDEN_PATH)
    keys = sorted({k.rsplit("_", 1)[0] for k in g.files if k.endswith("_geom")})
    out = []
    for key in keys:
        geom = g[key + "_geom"]
        ndim, L = int(geom[0]), int(geom[1])
        Kd = tuple(int(v) for v in geom[2:2 + ndim])
        Jd = tuple(int(v) for v in geom[2 + ndim:2 + 2 * ndim])
        hs = [g[key + "_h%d" % d] for d in range(ndim)]
        out.append(dict(key=key, ndim=ndim, L=L, Kd=Kd, Jd=Jd, hs=hs, tm=g[key + "_tm"], ck=g[key + "_ck"],
                        fm=g[key + "_fm"], fwd=g[key + "_fwd"], adj=g[key + "_adj"],
                        complex_kernel=np.iscomplexobj(hs[0])))
    return out
