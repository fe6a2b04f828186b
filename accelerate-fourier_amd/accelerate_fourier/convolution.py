"""Karatsuba convolution (Data.Array.Accelerate.Convolution.Adhoc).

``karatsuba(x, y)`` is ``karatsuba :: Transform2 (sh :. Int) a`` for Int64
(Convolution/Adhoc.hs:36-64): equal-length rows, result length 2L-1 (0 for
L = 0), wrap-around Int64 arithmetic, bit-exact on the GPU.
``cyclic(karatsuba)`` is ``Convolution.cyclic karatsuba`` (Adhoc.hs:104-111).
``complex(karatsuba)`` and ``complex(cyclic(karatsuba))`` are
``Convolution.complex`` over them (Adhoc.hs:114-133): complex Int64 operands
as (re, im) pairs of int64 tensors.
"""
from __future__ import annotations

import ctypes

import torch

from . import _abi, _arrays


def _check_pair(x, y):
    for t, nm in ((x, "x"), (y, "y")):
        _arrays._check_tensor(t, nm)
        if t.dtype != torch.int64:
            raise TypeError(f"{nm} must be int64")
    if x.shape != y.shape:
        # Adhoc.hs:33-34: both arrays must have the same size
        raise ValueError("Karatsuba operands must have the same shape")


def karatsuba(x, y, out=None):
    _check_pair(x, y)
    L = x.shape[-1]
    batch = _arrays.batch_of(tuple(x.shape))
    zl = 2 * L - 1 if L > 0 else 0
    if out is None:
        out = torch.empty(tuple(x.shape[:-1]) + (zl,), dtype=torch.int64, device=x.device)
    _arrays.check_real_out(out, tuple(x.shape[:-1]) + (zl,), torch.int64)
    _abi.check(_abi.lib().af_karatsuba_i64(batch, L, x.data_ptr(), y.data_ptr(), out.data_ptr(),
                                           _arrays.stream_handle()), "af_karatsuba_i64")
    return out


def karatsuba_ragged(offsets, x, y, out=None):
    """Pairs of uneven lengths packed back to back; pair p occupies
    offsets[p]:offsets[p+1] of x and y and 2*offsets[p]-p : +2L-1 of the result."""
    offsets = [int(o) for o in offsets]
    npairs = len(offsets) - 1
    if npairs < 0 or offsets[0] != 0:
        raise ValueError("offsets start at 0 and hold npairs + 1 entries")
    if any(b - a < 1 for a, b in zip(offsets, offsets[1:])):
        raise ValueError("every pair needs a length of at least 1 (offsets strictly increasing)")
    for t, nm in ((x, "x"), (y, "y")):
        _arrays._check_tensor(t, nm)
        if t.dtype != torch.int64 or t.dim() != 1:
            raise TypeError(f"{nm} must be a flat int64 tensor")
        if t.numel() < offsets[-1]:
            raise ValueError(f"{nm} holds {t.numel()} coefficients, the offsets need {offsets[-1]}")
    zn = 2 * offsets[-1] - npairs if npairs > 0 else 0
    if out is None:
        out = torch.empty(zn, dtype=torch.int64, device=x.device)
    _arrays._check_written(out, "out")
    if out.dtype != torch.int64 or out.dim() != 1 or out.numel() < zn:
        raise ValueError(f"out must be a flat int64 tensor of at least {zn} entries")
    arr = _abi.int64_array(offsets)
    _abi.check(_abi.lib().af_karatsuba_ragged_i64(npairs, ctypes.cast(arr, ctypes.POINTER(ctypes.c_int64)),
                                                  x.data_ptr(), y.data_ptr(), out.data_ptr(),
                                                  _arrays.stream_handle()), "af_karatsuba_ragged_i64")
    return out


def cyclic(conv):
    """Convolution.cyclic: wrap the tail of a linear convolution (Adhoc.hs:104-111)."""
    if conv is karatsuba:
        return _cyclic_karatsuba
    raise NotImplementedError("cyclic wraps the GPU Karatsuba only")


def complex(conv):
    """Convolution.complex (Adhoc.hs:114-133): a real convolution made complex
    by the Gauss form, re = xr*yr - xi*yi, im = (xr+xi)*(yr+yi) - xr*yr - xi*yi.
    The operands and the result are (re, im) pairs of int64 tensors; wrap-around
    arithmetic, bit-exact."""
    if conv is karatsuba:
        cyc = 0
    elif conv is _cyclic_karatsuba:
        cyc = 1
    else:
        raise NotImplementedError("complex wraps the GPU Karatsuba and its cyclic form only")

    def cconv(x, y, out=None):
        if not (isinstance(x, tuple) and isinstance(y, tuple) and len(x) == 2 and len(y) == 2):
            raise TypeError("complex Int64 operands are (re, im) pairs of int64 tensors")
        _check_pair(x[0], y[0])
        _check_pair(x[1], y[1])
        if x[0].shape != x[1].shape:
            raise ValueError("real and imaginary parts must have the same shape")
        L = x[0].shape[-1]
        zl = L if cyc else (2 * L - 1 if L > 0 else 0)
        shape = tuple(x[0].shape[:-1]) + (zl,)
        if out is None:
            out = (torch.empty(shape, dtype=torch.int64, device=x[0].device),
                   torch.empty(shape, dtype=torch.int64, device=x[0].device))
        if not (isinstance(out, tuple) and len(out) == 2):
            raise TypeError("the complex result is an (re, im) pair of int64 tensors")
        _arrays.check_real_out(out[0], shape, torch.int64, "out[0]")
        _arrays.check_real_out(out[1], shape, torch.int64, "out[1]")
        _abi.check(_abi.lib().af_karatsuba_complex_i64(
            _arrays.batch_of(tuple(x[0].shape)), L, cyc, x[0].data_ptr(), x[1].data_ptr(), y[0].data_ptr(),
            y[1].data_ptr(), out[0].data_ptr(), out[1].data_ptr(), _arrays.stream_handle()),
            "af_karatsuba_complex_i64")
        return out
    return cconv


def _cyclic_karatsuba(x, y, out=None):
    _check_pair(x, y)
    L = x.shape[-1]
    if out is None:
        out = torch.empty_like(x)
    _arrays.check_real_out(out, x.shape, torch.int64)
    _abi.check(_abi.lib().af_karatsuba_cyclic_i64(_arrays.batch_of(tuple(x.shape)), L, x.data_ptr(),
                                                  y.data_ptr(), out.data_ptr(), _arrays.stream_handle()),
               "af_karatsuba_cyclic_i64")
    return out
