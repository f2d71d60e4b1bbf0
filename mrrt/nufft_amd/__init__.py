"""MI355X-native table1 NUFFT (drop-in for mrrt.nufft's table and sparse modes).

Public names mirror ``mrrt.nufft`` (mrrt/nufft/__init__.py:1-15):
``NufftBase``, ``nufft_forward``, ``nufft_adj``, ``kaiser_bessel``,
``kaiser_bessel_ft``, ``BeattyKernel`` parameters via ``beatty_alpha``.
Beyond them: ``batched.NufftFrames`` (frames x coils), ``autograd``,
``toeplitz.ToeplitzNormal``, ``mri.SenseOperator`` (multi-coil encoding)
and ``autotune``.
"""

from ._kernel_math import beatty_alpha, kaiser_bessel, kaiser_bessel_ft  # noqa: F401
from ._nufft import NufftBase, nufft_adj, nufft_forward  # noqa: F401

__version__ = "0.1.0"

__all__ = [
    "NufftBase",
    "nufft_forward",
    "nufft_adj",
    "beatty_alpha",
    "kaiser_bessel",
    "kaiser_bessel_ft",
]
