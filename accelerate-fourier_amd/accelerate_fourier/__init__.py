"""accelerate-fourier's transform path on MI355X (gfx950).

Host-side mirror of the reference's public functions over the C ABI:

  Data.Array.Accelerate.Fourier.Planned   -> accelerate_fourier.planned
  Data.Array.Accelerate.Fourier.Real      -> accelerate_fourier.real
  Data.Array.Accelerate.Convolution.Adhoc -> accelerate_fourier.convolution
  Data.Array.Accelerate.Fourier.Sign      -> accelerate_fourier.sign
  Data.Array.Accelerate.Fourier.Utility   -> accelerate_fourier.utility

Arrays are torch tensors resident on the GPU (the shape ``sh :. n`` is the
tensor shape, transform axis last).  Complex arrays are either complex
tensors (interleaved) or ``(re, im)`` pairs of real tensors (planar, as
Accelerate stores them).
"""
from . import _abi, convolution, planned, real, sign, utility  # noqa: F401
from .planned import Transform, convolve_cyclic, transform  # noqa: F401
from .sign import Sign, forward, inverse  # noqa: F401
