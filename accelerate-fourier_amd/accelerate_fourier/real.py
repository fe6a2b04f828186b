"""Real-signal transforms (Data.Array.Accelerate.Fourier.Real).

``to_spectrum(sub, x)`` is ``toSpectrum :: Transform (sh:.Int) (Complex a) ->
Acc (Array (sh:.Int) a) -> Acc (Array (sh:.Int) (Complex a))`` (Real.hs:49-70)
with ``sub`` the half-length complex transform; the result is the full (not
halved) spectrum.  ``from_spectrum(sub, s)`` is ``fromSpectrum`` (Real.hs:92-107),
unnormalised.  Both run as one fused gfx950 pass: the half-length FFT with the
(de)interleaving and the (un)tangling folded into its loads and stores.

``two_to_spectrum(sub, x, y)`` / ``two_from_spectrum(sub, a, b)`` are
``twoToSpectrum`` / ``twoFromSpectrum`` (Real.hs:129-196): two real signals
through one complex transform ``sub`` of their full length, with
``untangle_spectra`` / ``entangle_spectra`` (Real.hs:168-236) as the packing
steps.  The reference's arrays of pairs are passed as their two components.
"""
from __future__ import annotations

import torch

from . import _abi, _arrays
from .planned import Transform, get_plan


def _real_prec(x):
    if x.dtype == torch.float64:
        return _abi.F64
    if x.dtype == torch.float32:
        return _abi.F32
    raise TypeError("real arrays are float64 or float32")


def _check_real(x, what):
    _arrays._check_tensor(x, what)
    if x.is_complex():
        raise TypeError(f"{what} must be real")


def to_spectrum(sub: Transform, x, out=None, planar=False):
    x = _arrays.materialize(x)
    _check_real(x, "x")
    n = x.shape[-1]
    if n % 2:
        raise ValueError("toSpectrum needs an even length (Real.hs:45-47)")
    if sub.n != n // 2:
        raise ValueError("the sub-transform must have half the signal length")
    prec = _real_prec(x)
    if out is None:
        if planar:
            out = (torch.empty_like(x), torch.empty_like(x))
        else:
            out = torch.empty(x.shape, dtype=_arrays.complex_dtype(prec), device=x.device)
    _arrays.check_complex_out(out, x.shape, prec)
    ore, oim = _arrays.ptrs(out)
    p = get_plan(n, sub.sign.value, _abi.R2C_FULL, prec, flags=sub.flags)
    _abi.check(_abi.lib().af_to_spectrum(p.handle, _arrays.batch_of(tuple(x.shape)), x.data_ptr(), ore, oim, None,
                                         _arrays.stream_handle()), "af_to_spectrum")
    return out


def from_spectrum(sub: Transform, spec, out=None):
    spec = _arrays.materialize(spec)
    re, im, shape = _arrays.complex_view(spec, "spectrum")
    n = shape[-1]
    if n % 2:
        raise ValueError("fromSpectrum needs an even length (Real.hs:88-90)")
    if sub.n != n // 2:
        raise ValueError("the sub-transform must have half the signal length")
    prec = _arrays.precision_of_complex(spec)
    if out is None:
        dev = spec[0].device if isinstance(spec, tuple) else spec.device
        out = torch.empty(shape, dtype=_arrays.real_dtype(prec), device=dev)
    _arrays.check_real_out(out, shape, _arrays.real_dtype(prec))
    p = get_plan(n, sub.sign.value, _abi.C2R_FULL, prec, flags=sub.flags)
    _abi.check(_abi.lib().af_from_spectrum(p.handle, _arrays.batch_of(shape), re, im, out.data_ptr(), None,
                                           _arrays.stream_handle()), "af_from_spectrum")
    return out


def vjp_to_spectrum(g, out=None):
    """Adjoint of to_spectrum(transform(forward, n/2), .) applied to a complex cotangent."""
    g = _arrays.materialize(g)
    re, im, shape = _arrays.complex_view(g, "cotangent")
    n = shape[-1]
    if n % 2:
        raise ValueError("toSpectrum needs an even length (Real.hs:45-47)")
    prec = _arrays.precision_of_complex(g)
    if out is None:
        dev = g[0].device if isinstance(g, tuple) else g.device
        out = torch.empty(shape, dtype=_arrays.real_dtype(prec), device=dev)
    _arrays.check_real_out(out, shape, _arrays.real_dtype(prec))
    p = get_plan(n, -1, _abi.R2C_FULL, prec)
    _abi.check(_abi.lib().af_vjp_to_spectrum(p.handle, _arrays.batch_of(shape), re, im, out.data_ptr(), None,
                                             _arrays.stream_handle()), "af_vjp_to_spectrum")
    return out


def vjp_from_spectrum(g, out=None):
    """Adjoint of from_spectrum(transform(inverse, n/2), .) applied to a real cotangent."""
    g = _arrays.materialize(g)
    _check_real(g, "cotangent")
    n = g.shape[-1]
    if n % 2:
        raise ValueError("fromSpectrum needs an even length (Real.hs:88-90)")
    prec = _real_prec(g)
    if out is None:
        out = torch.empty(g.shape, dtype=_arrays.complex_dtype(prec), device=g.device)
    _arrays.check_complex_out(out, g.shape, prec)
    ore, oim = _arrays.ptrs(out)
    p = get_plan(n, 1, _abi.C2R_FULL, prec)
    _abi.check(_abi.lib().af_vjp_from_spectrum(p.handle, _arrays.batch_of(tuple(g.shape)), g.data_ptr(), ore, oim,
                                               None, _arrays.stream_handle()), "af_vjp_from_spectrum")
    return out


def convolve(x, k, out=None):
    """fromSpectrum (transform inverse (n/2)) (toSpectrum (transform forward (n/2)) x *
    toSpectrum (transform forward (n/2)) k): n times the cyclic convolution, one fused launch."""
    x, k = _arrays.materialize(x), _arrays.materialize(k)
    _check_real(x, "x")
    _check_real(k, "k")
    if x.shape != k.shape or x.dtype != k.dtype:
        raise ValueError("signal and kernel must have equal shape and dtype")
    n = x.shape[-1]
    prec = _real_prec(x)
    if out is None:
        out = torch.empty_like(x)
    _arrays.check_real_out(out, x.shape, x.dtype)
    p = get_plan(n, -1, _abi.REAL_CONV, prec)
    _abi.check(_abi.lib().af_real_convolve(p.handle, _arrays.batch_of(tuple(x.shape)), x.data_ptr(), k.data_ptr(),
                                           out.data_ptr(), _arrays.stream_handle()), "af_real_convolve")
    return out


def convolve_vjp(x, k, w, dx=None, dk=None):
    """Gradients of sum(w * convolve(x, k)) with respect to x and k (one fused launch)."""
    x, k, w = _arrays.materialize(x), _arrays.materialize(k), _arrays.materialize(w)
    for t, nm in ((x, "x"), (k, "k"), (w, "w")):
        _check_real(t, nm)
    # the fused adjoint indexes k, w, dx and dk by x's rows and length
    _arrays.check_same(k, x, "k", "x")
    _arrays.check_same(w, x, "w", "x")
    n = x.shape[-1]
    prec = _real_prec(x)
    dx = torch.empty_like(x) if dx is None else dx
    dk = torch.empty_like(k) if dk is None else dk
    _arrays.check_real_out(dx, x.shape, x.dtype, "dx")
    _arrays.check_real_out(dk, x.shape, x.dtype, "dk")
    p = get_plan(n, -1, _abi.REAL_CONV, prec)
    _abi.check(_abi.lib().af_vjp_real_convolve(p.handle, _arrays.batch_of(tuple(x.shape)), x.data_ptr(),
                                               k.data_ptr(), w.data_ptr(), dx.data_ptr(), dk.data_ptr(),
                                               _arrays.stream_handle()), "af_vjp_real_convolve")
    return dx, dk


def _check_pair_out(out, shape, prec):
    if not (isinstance(out, tuple) and len(out) == 2):
        raise TypeError("out is a pair of complex arrays (a, b)")
    _arrays.check_complex_out(out[0], shape, prec, "out[0]")
    _arrays.check_complex_out(out[1], shape, prec, "out[1]")


def _check_real_pair_out(out, shape, prec):
    if not (isinstance(out, tuple) and len(out) == 2):
        raise TypeError("out is a pair of real arrays (x, y)")
    _arrays.check_real_out(out[0], shape, _arrays.real_dtype(prec), "out[0]")
    _arrays.check_real_out(out[1], shape, _arrays.real_dtype(prec), "out[1]")


def _two_shape(a, b, what):
    if tuple(a.shape) != tuple(b.shape):
        raise ValueError(f"the two {what} must have equal shapes")


def two_to_spectrum(sub: Transform, x, y, out=None):
    """Real.twoToSpectrum sub (Real.hs:129-135): the spectra (a, b) of the real
    rows x and y from one complex transform of x + i y (any length)."""
    x, y = _arrays.materialize(x), _arrays.materialize(y)
    _check_real(x, "x")
    _check_real(y, "y")
    _two_shape(x, y, "signals")
    if x.dtype != y.dtype:
        raise TypeError("the two signals must have the same dtype")
    n = x.shape[-1]
    if sub.n != n:
        raise ValueError("the sub-transform must have the signals' length")
    prec = _real_prec(x)
    if out is None:
        ct = _arrays.complex_dtype(prec)
        out = (torch.empty(x.shape, dtype=ct, device=x.device), torch.empty(x.shape, dtype=ct, device=x.device))
    _check_pair_out(out, x.shape, prec)
    are, aim = _arrays.ptrs(out[0])
    bre, bim = _arrays.ptrs(out[1])
    p = sub.plan(prec)
    _abi.check(_abi.lib().af_two_to_spectrum(p.handle, _arrays.batch_of(tuple(x.shape)), x.data_ptr(), y.data_ptr(),
                                             are, aim, bre, bim, _arrays.stream_handle()), "af_two_to_spectrum")
    return out


def two_from_spectrum(sub: Transform, a, b, out=None):
    """Real.twoFromSpectrum sub (Real.hs:190-196): the real and imaginary
    parts (x, y) of sub applied to the entangled spectrum a + i b."""
    a, b = _arrays.materialize(a), _arrays.materialize(b)
    are, aim, shape = _arrays.complex_view(a, "a")
    bre, bim, shape_b = _arrays.complex_view(b, "b")
    if tuple(shape) != tuple(shape_b):
        raise ValueError("the two spectra must have equal shapes")
    n = shape[-1]
    if sub.n != n:
        raise ValueError("the sub-transform must have the spectra's length")
    prec = _arrays.precision_of_complex(a)
    if prec != _arrays.precision_of_complex(b):
        raise TypeError("the two spectra must have the same precision")
    if out is None:
        dev = a[0].device if isinstance(a, tuple) else a.device
        rt = _arrays.real_dtype(prec)
        out = (torch.empty(shape, dtype=rt, device=dev), torch.empty(shape, dtype=rt, device=dev))
    _check_real_pair_out(out, shape, prec)
    p = sub.plan(prec)
    _abi.check(_abi.lib().af_two_from_spectrum(p.handle, _arrays.batch_of(shape), are, aim, bre, bim,
                                               out[0].data_ptr(), out[1].data_ptr(), _arrays.stream_handle()),
               "af_two_from_spectrum")
    return out


def untangle_spectra(z, out=None):
    """Real.untangleSpectra (Real.hs:168-187): (a_k, b_k) = ((z_k + conj z_-k) / 2,
    (z_k - conj z_-k) (-i / 2)) along the last axis."""
    z = _arrays.materialize(z)
    zre, zim, shape = _arrays.complex_view(z, "z")
    prec = _arrays.precision_of_complex(z)
    if out is None:
        out = (_arrays.empty_like_complex(z), _arrays.empty_like_complex(z))
    _check_pair_out(out, shape, prec)
    are, aim = _arrays.ptrs(out[0])
    bre, bim = _arrays.ptrs(out[1])
    _abi.check(_abi.lib().af_untangle_spectra(prec, _arrays.batch_of(shape), shape[-1], zre, zim, are, aim, bre, bim,
                                              _arrays.stream_handle()), "af_untangle_spectra")
    return out


def entangle_spectra(c, d, out=None):
    """Real.entangleSpectra (Real.hs:208-236): c + d i."""
    c, d = _arrays.materialize(c), _arrays.materialize(d)
    cre, cim, shape = _arrays.complex_view(c, "c")
    dre, dim, shape_d = _arrays.complex_view(d, "d")
    if tuple(shape) != tuple(shape_d):
        raise ValueError("the two spectra must have equal shapes")
    prec = _arrays.precision_of_complex(c)
    if _arrays.precision_of_complex(d) != prec:
        raise TypeError("the two spectra must have the same precision")
    if out is None:
        out = _arrays.empty_like_complex(c)
    _arrays.check_complex_out(out, shape, prec)
    zre, zim = _arrays.ptrs(out)
    _abi.check(_abi.lib().af_entangle_spectra(prec, _arrays.batch_of(shape), shape[-1], cre, cim, dre, dim, zre, zim,
                                              _arrays.stream_handle()), "af_entangle_spectra")
    return out


def two_to_spectrum_2d(t0: Transform, t1: Transform, x, y, out=None):
    """Real.twoToSpectrum2d (transform2d t0 t1) (Real.hs:138-145) over (..., H, W)."""
    x, y = _arrays.materialize(x), _arrays.materialize(y)
    _check_real(x, "x")
    _check_real(y, "y")
    _two_shape(x, y, "signals")
    if x.dim() < 2 or x.shape[-1] != t0.n or x.shape[-2] != t1.n:
        raise ValueError("the sub-transforms must match the last two axes")
    prec = _real_prec(x)
    if out is None:
        ct = _arrays.complex_dtype(prec)
        out = (torch.empty(x.shape, dtype=ct, device=x.device), torch.empty(x.shape, dtype=ct, device=x.device))
    _check_pair_out(out, x.shape, prec)
    are, aim = _arrays.ptrs(out[0])
    bre, bim = _arrays.ptrs(out[1])
    _abi.check(_abi.lib().af_two_to_spectrum_2d(t0.plan(prec).handle, t1.plan(prec).handle,
                                                _arrays.batch_of(tuple(x.shape)[:-1]), x.data_ptr(), y.data_ptr(),
                                                are, aim, bre, bim, _arrays.stream_handle()), "af_two_to_spectrum_2d")
    return out


def two_from_spectrum_2d(t0: Transform, t1: Transform, a, b, out=None):
    """Real.twoFromSpectrum2d (transform2d t0 t1) (Real.hs:199-206)."""
    a, b = _arrays.materialize(a), _arrays.materialize(b)
    are, aim, shape = _arrays.complex_view(a, "a")
    bre, bim, shape_b = _arrays.complex_view(b, "b")
    if tuple(shape) != tuple(shape_b):
        raise ValueError("the two spectra must have equal shapes")
    if len(shape) < 2 or shape[-1] != t0.n or shape[-2] != t1.n:
        raise ValueError("the sub-transforms must match the last two axes")
    prec = _arrays.precision_of_complex(a)
    if out is None:
        dev = a[0].device if isinstance(a, tuple) else a.device
        rt = _arrays.real_dtype(prec)
        out = (torch.empty(shape, dtype=rt, device=dev), torch.empty(shape, dtype=rt, device=dev))
    _check_real_pair_out(out, shape, prec)
    _abi.check(_abi.lib().af_two_from_spectrum_2d(t0.plan(prec).handle, t1.plan(prec).handle,
                                                  _arrays.batch_of(shape[:-1]), are, aim, bre, bim,
                                                  out[0].data_ptr(), out[1].data_ptr(), _arrays.stream_handle()),
               "af_two_from_spectrum_2d")
    return out


def untangle_spectra_2d(z, out=None):
    """Real.untangleSpectra2d (Real.hs:175-180): partners at Cyclic.reverse2d."""
    z = _arrays.materialize(z)
    zre, zim, shape = _arrays.complex_view(z, "z")
    if len(shape) < 2:
        raise ValueError("untangleSpectra2d needs at least two axes")
    prec = _arrays.precision_of_complex(z)
    if out is None:
        out = (_arrays.empty_like_complex(z), _arrays.empty_like_complex(z))
    _check_pair_out(out, shape, prec)
    are, aim = _arrays.ptrs(out[0])
    bre, bim = _arrays.ptrs(out[1])
    _abi.check(_abi.lib().af_untangle_spectra_2d(prec, _arrays.batch_of(shape[:-1]), shape[-2], shape[-1], zre, zim,
                                                 are, aim, bre, bim, _arrays.stream_handle()),
               "af_untangle_spectra_2d")
    return out
