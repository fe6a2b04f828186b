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
