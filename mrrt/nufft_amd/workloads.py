"""The five benchmark configurations of the table1 NUFFT (seeded inputs).

Each function returns a dict with the operator geometry (``Nd``, ``Kd``,
``Jd``, ``Ld``, ``precision``), the sampling points ``om`` (float64, radians in
[-pi, pi)) and the seeded inputs of that configuration.  The same inputs are
used by the parity tests and by the benchmark, so a benchmark figure always
refers to a configuration whose results are checked.

c1  2-D radial, Nd 128^2, Kd 256^2, J 6, L 1024, 64 spokes x 256 readout,
    complex64 (the size the reference runs on a CPU in seconds)
c2  2-D golden-angle radial, Nd 256^2, Kd 512^2, J 6, L 1024,
    402 spokes x 512 readout (205,824 points); image = complex Gaussian +
    10 random ellipses; complex64 or complex128
c3  3-D kooshball, Nd 128^3, Kd 256^3, J 6, L 512, 8,192 spokes x 256
    readout (2,097,152 points) on a seeded Fibonacci sphere; complex64
c4  24 frames x 32 coils, golden-angle frames of 13..89 spokes x 512
    readout (per-frame spoke counts seeded), Nd 256^2, Kd 512^2, J 6,
    complex64, with smooth random complex coil maps and per-frame images
c5  variable-density spiral, 16 interleaves x 4,096 samples (seeded
    jitter), Nd 256^2, Kd 512^2, J 6, 8 coils, complex64; target samples
    y = A(x, om_pert) from a perturbed trajectory
"""

import numpy as np

from . import trajectories as trj

__all__ = ["c1", "c2", "c3", "c4", "c5"]


def c1():
    Nd = (128, 128)
    return dict(name="c1", Nd=Nd, Kd=(256, 256), Jd=6, Ld=1024, precision="single",
                om=trj.radial_2d(64, 256), x=trj.complex_gaussian(Nd, seed=0))


def c2(precision="single"):
    Nd = (256, 256)
    x = trj.complex_gaussian(Nd, seed=0) + trj.ellipse_phantom(Nd, 10, seed=0)
    return dict(name="c2", Nd=Nd, Kd=(512, 512), Jd=6, Ld=1024, precision=precision,
                om=trj.golden_angle_2d(402, 512), x=x)


def c3():
    Nd = (128, 128, 128)
    return dict(name="c3", Nd=Nd, Kd=(256, 256, 256), Jd=6, Ld=512, precision="single",
                om=trj.kooshball_3d(8192, 256, seed=0), x=trj.complex_gaussian(Nd, seed=0))


def c4(nframes=24, ncoils=32):
    """Per-frame trajectories (list of arrays), coil maps (C,) + Nd and one
    image per frame (F,) + Nd, all seeded."""
    Nd = (256, 256)
    frames = trj.golden_angle_frames(nframes, 512, 13, 89, seed=0)
    return dict(name="c4", Nd=Nd, Kd=(512, 512), Jd=6, Ld=1024, precision="single", frames=frames,
                ncoils=ncoils, smaps=trj.coil_maps(ncoils, Nd, seed=1),
                x=trj.complex_gaussian((nframes,) + Nd, seed=2))


def c5(ncoils=8):
    """Spiral points, coil images x (C,) + Nd and the perturbed trajectory
    the target samples y = A(x, om_pert) come from."""
    Nd = (256, 256)
    om = trj.spiral_2d(16, 4096, seed=0)
    rng = np.random.default_rng(4)
    om_pert = np.clip(om + 1e-3 * rng.standard_normal(om.shape), -np.pi, np.pi - 1e-9)
    return dict(name="c5", Nd=Nd, Kd=(512, 512), Jd=6, Ld=1024, precision="single", om=om, om_pert=om_pert,
                ncoils=ncoils, x=trj.complex_gaussian((ncoils,) + Nd, seed=5))
