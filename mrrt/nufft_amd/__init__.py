"""MI355X-native table1 NUFFT (drop-in for mrrt.nufft's table and sparse modes).

Public names mirror ``mrrt.nufft`` (mrrt/nufft/__init__.py:1-15):
``dtft``, ``dtft_adj`` (exact NUDFT, on the GPU), ``BeattyKernel``,
``KaiserBesselKernel``, ``NufftBase``, ``kaiser_bessel``, ``kaiser_bessel_ft``,
plus ``nufft_forward`` / ``nufft_adj`` and ``beatty_alpha``.  Beyond them:
``batched.NufftFrames`` (frames x coils), ``autograd``,
``toeplitz.ToeplitzNormal``, ``mri.SenseOperator`` (multi-coil encoding),
``workloads`` (the benchmark configurations) and ``autotune``.
"""

from ._dtft import dtft, dtft_adj  # noqa: F401
from ._kernel_math import beatty_alpha, kaiser_bessel, kaiser_bessel_ft  # noqa: F401
from ._kernels import BeattyKernel, KaiserBesselKernel  # noqa: F401
from ._nufft import NufftBase, nufft_adj, nufft_forward  # noqa: F401

__version__ = "0.2.0"

__all__ = [
    "dtft",
    "dtft_adj",
    "BeattyKernel",
    "KaiserBesselKernel",
    "NufftBase",
    "kaiser_bessel",
    "kaiser_bessel_ft",
    "nufft_forward",
    "nufft_adj",
    "beatty_alpha",
]
