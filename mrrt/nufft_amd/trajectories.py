"""Synthetic k-space trajectories and test images (seeded, float64).

Omega values are in radians; each axis is sampled in [-pi, pi).
These produce the workloads the operator is measured and checked on:
radial, golden-angle radial, 3-D kooshball, variable-density spiral,
and multi-frame golden-angle with uneven spoke counts.
"""

import numpy as np

__all__ = [
    "radial_2d",
    "golden_angle_2d",
    "kooshball_3d",
    "spiral_2d",
    "golden_angle_frames",
    "ellipse_phantom",
    "complex_gaussian",
    "coil_maps",
]

GOLDEN = np.pi * (np.sqrt(5.0) - 1.0) / 2.0


def _readout(nread):
    return -np.pi + 2 * np.pi * np.arange(nread) / nread


def radial_2d(nspokes, nread):
    """Uniform radial: angles pi*s/nspokes, readout r_j = -pi + 2 pi j / nread."""
    th = np.pi * np.arange(nspokes) / nspokes
    r = _readout(nread)
    return np.stack([(r[None, :] * np.cos(th[:, None])).ravel(),
                     (r[None, :] * np.sin(th[:, None])).ravel()], axis=-1)


def golden_angle_2d(nspokes, nread, first=0):
    """Golden-angle radial: angle of spoke s is (first + s) * pi (sqrt5-1)/2 mod pi."""
    th = np.mod((first + np.arange(nspokes)) * GOLDEN, np.pi)
    r = _readout(nread)
    return np.stack([(r[None, :] * np.cos(th[:, None])).ravel(),
                     (r[None, :] * np.sin(th[:, None])).ravel()], axis=-1)


def kooshball_3d(nspokes, nread, seed=0):
    """3-D radial: spoke directions on a Fibonacci sphere, randomly rotated."""
    rng = np.random.default_rng(seed)
    i = np.arange(nspokes) + 0.5
    z = 1 - 2 * i / nspokes
    phi = np.pi * (1 + 5 ** 0.5) * i
    rxy = np.sqrt(1 - z * z)
    dirs = np.stack([rxy * np.cos(phi), rxy * np.sin(phi), z], axis=-1)
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    dirs = dirs @ q.T
    r = _readout(nread)
    om = r[None, :, None] * dirs[:, None, :]
    return om.reshape(-1, 3)


def spiral_2d(ninterleaves, nsamples, seed=0, jitter=1e-3, density_power=1.5):
    """Variable-density Archimedean-like spiral with seeded jitter, |omega| < pi."""
    rng = np.random.default_rng(seed)
    s = np.linspace(0, 1, nsamples, endpoint=False)
    r = np.pi * 0.999 * s ** density_power
    turns = 16.0
    out = []
    for k in range(ninterleaves):
        th = 2 * np.pi * turns * s + 2 * np.pi * k / ninterleaves
        om = np.stack([r * np.cos(th), r * np.sin(th)], axis=-1)
        out.append(om)
    om = np.concatenate(out, axis=0)
    om = om + jitter * rng.standard_normal(om.shape)
    return np.clip(om, -np.pi, np.pi - 1e-9)


def golden_angle_frames(nframes, nread, smin=13, smax=89, seed=0):
    """Per-frame golden-angle radial trajectories with uneven spoke counts.

    Spoke counts S_f ~ U{smin..smax}; the golden-angle index continues across
    frames.  Returns a list of (S_f*nread, 2) arrays.
    """
    rng = np.random.default_rng(seed)
    counts = rng.integers(smin, smax + 1, size=nframes)
    frames, first = [], 0
    for c in counts:
        frames.append(golden_angle_2d(int(c), nread, first=first))
        first += int(c)
    return frames


def complex_gaussian(shape, seed=0, dtype=np.complex128):
    rng = np.random.default_rng(seed)
    return ((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)).astype(dtype)


def ellipse_phantom(shape, nellipses=10, seed=0):
    """Sum of random complex-valued ellipses (2-D), float64 complex."""
    rng = np.random.default_rng(seed)
    ny, nx = shape
    y, x = np.meshgrid(np.linspace(-1, 1, ny), np.linspace(-1, 1, nx), indexing="ij")
    img = np.zeros(shape, np.complex128)
    for _ in range(nellipses):
        cx_, cy_ = rng.uniform(-0.5, 0.5, 2)
        a, b = rng.uniform(0.05, 0.5, 2)
        th = rng.uniform(0, np.pi)
        amp = rng.standard_normal() + 1j * rng.standard_normal()
        xr = (x - cx_) * np.cos(th) + (y - cy_) * np.sin(th)
        yr = -(x - cx_) * np.sin(th) + (y - cy_) * np.cos(th)
        img[(xr / a) ** 2 + (yr / b) ** 2 <= 1] += amp
    return img


def coil_maps(ncoils, shape, seed=0):
    """Smooth random complex coil sensitivities, (ncoils,) + shape complex128.

    Coil c is a Gaussian bump exp(-|r - r_c|^2 / w_c^2) times a linear phase
    exp(i (phi_c + g_c . r)) on the grid r in [-1, 1]^d, with the centre r_c,
    width w_c, phase offset phi_c and gradient g_c drawn from a seeded rng;
    the maps are normalised so that sum_c |S_c|^2 has maximum 1.
    """
    rng = np.random.default_rng(seed)
    shape = tuple(int(n) for n in shape)
    r = np.meshgrid(*[np.linspace(-1, 1, n) for n in shape], indexing="ij")
    maps = np.empty((ncoils,) + shape, np.complex128)
    for c in range(ncoils):
        cen = rng.uniform(-1.2, 1.2, len(shape))
        width = rng.uniform(0.6, 1.4)
        phi = rng.uniform(0, 2 * np.pi)
        grad = rng.uniform(-2, 2, len(shape))
        d2 = sum((r[d] - cen[d]) ** 2 for d in range(len(shape)))
        ph = phi + sum(grad[d] * r[d] for d in range(len(shape)))
        maps[c] = np.exp(-d2 / width ** 2) * np.exp(1j * ph)
    return maps / np.sqrt(np.max(np.sum(np.abs(maps) ** 2, axis=0)))
