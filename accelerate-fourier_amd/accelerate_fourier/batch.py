"""Uneven batches: many transforms of different lengths in one graph launch.

``UnevenBatch(lengths, sign, precision)`` applies ``Planned.transform sign
lengths[i]`` (Planned.hs:103-106) to each of ``len(lengths)`` rows packed back
to back in one flat complex array.  All rows run as grouped launches captured
in one hipGraph, rows with intermediates in cache-sized chunks.
"""
from __future__ import annotations

import ctypes

import torch

from . import _abi, _arrays
from .sign import Sign


class UnevenBatch:
    def __init__(self, lengths, sign: Sign, precision: str = "f32"):
        if not isinstance(sign, Sign):
            sign = Sign(int(sign))
        self.lengths = [int(n) for n in lengths]
        self.total = sum(self.lengths)
        self.prec = _abi.F64 if precision == "f64" else _abi.F32
        arr = _abi.int64_array(self.lengths)
        h = ctypes.c_void_p()
        _abi.check(_abi.lib().af_batch_create(len(self.lengths), ctypes.cast(arr, ctypes.POINTER(ctypes.c_int64)),
                                              sign.value, self.prec, ctypes.byref(h)), "af_batch_create")
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _abi is not None and _abi._lib is not None:
            try:
                _abi._lib.af_batch_destroy(h)
            except Exception:  # interpreter teardown
                pass
            self.handle = None

    def describe(self) -> str:
        """Plan counts per algorithm, launches, chunks and scratch on the first
        line, then one line per grouped launch: its chunk, phase, kernel key,
        row count and first row."""
        buf = ctypes.create_string_buffer(1 << 20)
        _abi.check(_abi.lib().af_batch_describe(self.handle, buf, len(buf)), "af_batch_describe")
        return buf.value.decode()

    def set_schedule(self, chunk_bytes: int = 0, lanes: int = 2) -> None:
        """af_batch_set_schedule: rows with intermediates in chunks of at most
        chunk_bytes of them (0: no chunking); lanes = 2 runs the rows without
        intermediates on a parallel branch of the graph, lanes = 3 (unchunked
        only) also the mixed-radix and Rader rows with intermediates on a third."""
        if int(chunk_bytes) < 0:
            raise ValueError("chunk_bytes must be >= 0")
        if int(lanes) not in (1, 2, 3):
            raise ValueError("lanes must be 1, 2 or 3")
        if int(lanes) == 3 and int(chunk_bytes) != 0:
            raise ValueError("three lanes need chunk_bytes = 0")
        _abi.check(_abi.lib().af_batch_set_schedule(self.handle, int(chunk_bytes), int(lanes)),
                   "af_batch_set_schedule")

    def tune(self) -> str:
        """af_batch_tune: times chunk sizes of the batch's graph on the current
        stream and keeps the fastest, then the radix order of each generic
        grouped launch; returns the report."""
        buf = ctypes.create_string_buffer(1 << 16)
        _abi.check(_abi.lib().af_batch_tune(self.handle, _arrays.stream_handle(), buf, len(buf)), "af_batch_tune")
        return buf.value.decode()

    def __call__(self, x, out=None):
        x = _arrays.materialize(x)
        re, im, shape = _arrays.complex_view(x)
        if len(shape) != 1 or shape[0] != self.total:
            raise ValueError(f"expected a flat array of {self.total} elements")
        if _arrays.precision_of_complex(x) != self.prec:
            raise TypeError("precision differs from the batch's")
        if out is None:
            out = _arrays.empty_like_complex(x)
        _arrays.check_complex_out(out, shape, self.prec)
        ore, oim = _arrays.ptrs(out)
        _abi.check(_abi.lib().af_batch_execute(self.handle, re, im, ore, oim, _arrays.stream_handle()),
                   "af_batch_execute")
        return out
