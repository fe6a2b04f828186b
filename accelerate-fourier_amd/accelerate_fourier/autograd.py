"""Thin torch.autograd.Function wrappers over the C ABI's vector-Jacobian
products, used for gradient checks (finite differences, gradcheck).  Forward
and backward both run the gfx950 kernels; torch only carries the tensors.

Gradient convention is torch's: for a complex-linear map y = A x the returned
gradient is A^H g.
"""
from __future__ import annotations

import torch

from . import planned, real
from .sign import forward, inverse


class TransformFn(torch.autograd.Function):
    """Planned.transform sign n with its adjoint (the opposite-sign transform)."""

    @staticmethod
    def forward(ctx, x, t):
        ctx.t = t
        return t(x.contiguous())

    @staticmethod
    def backward(ctx, g):
        return ctx.t.vjp(g.contiguous()), None


class ToSpectrumFn(torch.autograd.Function):
    """toSpectrum (transform forward (n/2)) on real rows."""

    @staticmethod
    def forward(ctx, x):
        return real.to_spectrum(planned.transform(forward, x.shape[-1] // 2), x.contiguous())

    @staticmethod
    def backward(ctx, g):
        return real.vjp_to_spectrum(g.contiguous())


class FromSpectrumFn(torch.autograd.Function):
    """fromSpectrum (transform inverse (n/2)) on full spectra."""

    @staticmethod
    def forward(ctx, s):
        return real.from_spectrum(planned.transform(inverse, s.shape[-1] // 2), s.contiguous())

    @staticmethod
    def backward(ctx, g):
        return real.vjp_from_spectrum(g.contiguous())


class RealConvFn(torch.autograd.Function):
    """y = fromSpectrum(toSpectrum x * toSpectrum k) with the fused adjoint kernel."""

    @staticmethod
    def forward(ctx, x, k):
        x = x.contiguous()
        k = k.contiguous()
        ctx.save_for_backward(x, k)
        return real.convolve(x, k)

    @staticmethod
    def backward(ctx, w):
        x, k = ctx.saved_tensors
        dx, dk = real.convolve_vjp(x, k, w.contiguous())
        return dx, dk
