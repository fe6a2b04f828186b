"""Normalisation (Data.Array.Accelerate.Fourier.Utility, Utility.hs:17-21)."""
from __future__ import annotations


def scale_down(x):
    """Multiply by recip n, n the inner extent (scaleDown)."""
    n = x[0].shape[-1] if isinstance(x, tuple) else x.shape[-1]
    if n == 0:  # recip 0 scales an empty array: nothing to do
        return x
    r = 1.0 / n
    if isinstance(x, tuple):
        return (x[0] * r, x[1] * r)
    return x * r
