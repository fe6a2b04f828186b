"""Normalisation (Data.Array.Accelerate.Fourier.Utility, Utility.hs:17-21)."""
from __future__ import annotations


def scale_down(x):
    """Multiply by recip n, n the inner extent (scaleDown)."""
    if isinstance(x, tuple):
        n = x[0].shape[-1]
        r = 1.0 / n
        return (x[0] * r, x[1] * r)
    n = x.shape[-1]
    return x * (1.0 / n)
