"""Array plumbing between torch tensors and the C ABI (no arithmetic here)."""
from __future__ import annotations

import torch

from . import _abi


def stream_handle():
    return torch.cuda.current_stream().cuda_stream


def precision_of_complex(x):
    if isinstance(x, tuple):
        dt = x[0].dtype
    else:
        dt = x.dtype
    if dt in (torch.complex128, torch.float64):
        return _abi.F64
    if dt in (torch.complex64, torch.float32):
        return _abi.F32
    raise TypeError(f"unsupported dtype {dt}: Complex Double / Complex Float only")


def real_dtype(prec):
    return torch.float64 if prec == _abi.F64 else torch.float32


def complex_dtype(prec):
    return torch.complex128 if prec == _abi.F64 else torch.complex64


def _check_tensor(t, what):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{what} must be a torch tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{what} must reside on the GPU")
    if not t.is_contiguous():
        raise ValueError(f"{what} must be contiguous (transform axis innermost)")


def materialize(x):
    """Resolve torch's lazy conjugate / negative views: the kernels read memory."""
    if isinstance(x, tuple):
        return tuple(materialize(t) for t in x)
    if isinstance(x, torch.Tensor):
        if x.is_conj():
            x = x.resolve_conj()
        if x.is_neg():
            x = x.resolve_neg()
    return x


def complex_view(x, what="input"):
    """(re_ptr, im_ptr or None, shape) of a complex array in either layout."""
    if isinstance(x, tuple):
        re, im = x
        _check_tensor(re, what + ".re")
        _check_tensor(im, what + ".im")
        if re.shape != im.shape or re.dtype != im.dtype:
            raise ValueError(f"{what}: planar components differ in shape or dtype")
        return re.data_ptr(), im.data_ptr(), tuple(re.shape)
    _check_tensor(x, what)
    if not x.is_complex():
        raise TypeError(f"{what} must be complex")
    return x.data_ptr(), None, tuple(x.shape)


def _check_written(t, what):
    """A tensor a kernel writes into: on the GPU, contiguous, no lazy view bit."""
    _check_tensor(t, what)
    if t.is_conj() or t.is_neg():
        raise ValueError(f"{what} is a lazily conjugated or negated view; pass a plain tensor")


def check_complex_out(out, shape, prec, what="out"):
    """A caller-supplied complex result: a planar (re, im) pair of real tensors
    or one complex tensor, of the given shape and precision, on the GPU and
    contiguous.  Raises instead of letting a kernel write past its end."""
    shape = tuple(shape)
    if isinstance(out, tuple):
        if len(out) != 2:
            raise ValueError(f"{what}: a planar result is a (re, im) pair")
        for t, nm in zip(out, (".re", ".im")):
            _check_written(t, what + nm)
            if t.is_complex() or t.dtype != real_dtype(prec):
                raise TypeError(f"{what + nm} must be {real_dtype(prec)}")
            if tuple(t.shape) != shape:
                raise ValueError(f"{what + nm} has shape {tuple(t.shape)}, expected {shape}")
        return out
    _check_written(out, what)
    if out.dtype != complex_dtype(prec):
        raise TypeError(f"{what} must be {complex_dtype(prec)}")
    if tuple(out.shape) != shape:
        raise ValueError(f"{what} has shape {tuple(out.shape)}, expected {shape}")
    return out


def check_real_out(out, shape, dtype, what="out"):
    """A caller-supplied real result of the given shape and dtype."""
    _check_written(out, what)
    if out.dtype != dtype:
        raise TypeError(f"{what} must be {dtype}")
    if tuple(out.shape) != tuple(shape):
        raise ValueError(f"{what} has shape {tuple(out.shape)}, expected {tuple(shape)}")
    return out


def check_same(t, ref, what, ref_what):
    """Operand t must match ref in shape and dtype (kernels index both by ref's extents)."""
    _check_tensor(t, what)
    if tuple(t.shape) != tuple(ref.shape) or t.dtype != ref.dtype:
        raise ValueError(f"{what} must have the shape and dtype of {ref_what} "
                         f"({tuple(ref.shape)}, {ref.dtype}); got ({tuple(t.shape)}, {t.dtype})")


def empty_like_complex(x, shape=None):
    if isinstance(x, tuple):
        s = shape or tuple(x[0].shape)
        return (torch.empty(s, dtype=x[0].dtype, device=x[0].device),
                torch.empty(s, dtype=x[0].dtype, device=x[0].device))
    return torch.empty(shape or tuple(x.shape), dtype=x.dtype, device=x.device)


def ptrs(x):
    if isinstance(x, tuple):
        return x[0].data_ptr(), x[1].data_ptr()
    return x.data_ptr(), None


def batch_of(shape):
    b = 1
    for s in shape[:-1]:
        b *= s
    return b
