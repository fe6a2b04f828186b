"""Planned transforms (Data.Array.Accelerate.Fourier.Planned).

``transform(sign, n)`` is the reference's ``transform :: Sign a -> Int ->
Transform (sh:.Int) (Complex a)`` (Planned.hs:103-106): a function over arrays
whose inner extent is n, unnormalised, sharing its plan (twiddles, chirps,
Rader kernels) between calls as the partially applied Haskell function shares
its cache.  The work runs in the gfx950 kernels behind the C ABI.
"""
from __future__ import annotations

import ctypes
import threading

from . import _abi, _arrays
from .sign import Sign

_plan_cache: dict = {}
_plan_lock = threading.Lock()


class Plan:
    """Owns one af_plan."""

    def __init__(self, length, sign, kind, precision, algorithm=_abi.ALG_AUTO, flags=0):
        self.desc = _abi.PlanDesc(length, sign, kind, precision, algorithm, flags)
        h = ctypes.c_void_p()
        _abi.check(_abi.lib().af_plan_create(ctypes.byref(self.desc), ctypes.byref(h)),
                   f"af_plan_create(n={length}, sign={sign}, kind={kind})")
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _abi is not None and _abi._lib is not None:
            try:
                _abi._lib.af_plan_destroy(h)
            except Exception:  # interpreter teardown
                pass
            self.handle = None

    def workspace_size(self, batch):
        n = ctypes.c_size_t()
        _abi.check(_abi.lib().af_plan_workspace_size(self.handle, batch, ctypes.byref(n)), "workspace")
        return n.value


def get_plan(length, sign, kind, precision, algorithm=_abi.ALG_AUTO, flags=0) -> Plan:
    key = (length, sign, kind, precision, algorithm, flags)
    with _plan_lock:
        p = _plan_cache.get(key)
        if p is None:
            p = Plan(length, sign, kind, precision, algorithm, flags)
            _plan_cache[key] = p
        return p


def describe(length, sign=-1, kind=_abi.C2C, precision=_abi.F32, algorithm=_abi.ALG_AUTO, flags=0) -> str:
    """Host-only text of the device plan for a length."""
    d = _abi.PlanDesc(length, sign, kind, precision, algorithm, flags)
    buf = ctypes.create_string_buffer(4096)
    st = _abi.lib().af_plan_describe(ctypes.byref(d), buf, len(buf))
    if st not in (_abi.SUCCESS, _abi.UNSUPPORTED):
        raise _abi.AfError(st, "af_plan_describe")
    return buf.value.decode()


class Transform:
    """``Planned.transform sign n`` as a callable over GPU arrays."""

    def __init__(self, sign: Sign, n: int, algorithm: int = _abi.ALG_AUTO, plain_twiddles: bool = False,
                 cooley_tukey: bool = False):
        if not isinstance(sign, Sign):
            sign = Sign(int(sign))
        if n < 0:
            raise ValueError("length must be non-negative")
        self.sign = sign
        self.n = int(n)
        self.algorithm = algorithm
        # cooley_tukey: single-pass lengths with several prime factors keep the
        # twiddled Cooley-Tukey order instead of the default prime-factor order
        self.flags = (_abi.FLAG_PLAIN_TWIDDLES if plain_twiddles else 0) | (_abi.FLAG_COOLEY_TUKEY if cooley_tukey else 0)

    def plan(self, precision) -> Plan:
        return get_plan(self.n, self.sign.value, _abi.C2C, precision, self.algorithm, self.flags)

    def __call__(self, x, out=None):
        x = _arrays.materialize(x)
        re, im, shape = _arrays.complex_view(x)
        if not shape or shape[-1] != self.n:
            # Planned.hs:98-101: the length must equal the inner extent
            raise ValueError(f"inner extent {shape[-1] if shape else None} differs from the planned length {self.n}")
        prec = _arrays.precision_of_complex(x)
        if out is None:
            out = _arrays.empty_like_complex(x)
        _arrays.check_complex_out(out, shape, prec)
        ore, oim = _arrays.ptrs(out)
        batch = _arrays.batch_of(shape)
        p = self.plan(prec)
        _abi.check(_abi.lib().af_execute(p.handle, batch, re, im, ore, oim, None, _arrays.stream_handle()),
                   "af_execute")
        return out

    def tune(self, batch, precision="f32", layout="planar"):
        """Autotune this transform's launches for `batch` rows in the given
        layout ("planar" or "interleaved"); later calls with that batch and
        layout use the fastest tile shapes.  Returns the tuner's report."""
        prec = _abi.F64 if precision in ("f64", _abi.F64) else _abi.F32
        if layout not in ("planar", "interleaved"):
            raise ValueError("layout must be 'planar' or 'interleaved'")
        p = self.plan(prec)
        buf = ctypes.create_string_buffer(8192)
        _abi.check(_abi.lib().af_plan_tune(p.handle, int(batch), 1 if layout == "planar" else 0,
                                           _arrays.stream_handle(), buf, len(buf)), "af_plan_tune")
        return buf.value.decode()

    def vjp(self, g, out=None):
        """Adjoint (conjugate transpose) applied to the cotangent g."""
        g = _arrays.materialize(g)
        re, im, shape = _arrays.complex_view(g, "cotangent")
        if shape[-1] != self.n:
            raise ValueError("cotangent has the wrong inner extent")
        prec = _arrays.precision_of_complex(g)
        if out is None:
            out = _arrays.empty_like_complex(g)
        _arrays.check_complex_out(out, shape, prec)
        ore, oim = _arrays.ptrs(out)
        p = self.plan(prec)
        _abi.check(_abi.lib().af_vjp_execute(p.handle, _arrays.batch_of(shape), re, im, ore, oim, None,
                                             _arrays.stream_handle()), "af_vjp_execute")
        return out


class Transform2d:
    """``Preprocessed.transform2d (SubTransformPair t0 t1)`` (Preprocessed.hs:141-148):
    t0 along the last axis, t1 along the one before, over arrays (..., H, W)."""

    def __init__(self, t0: Transform, t1: Transform):
        self.t0, self.t1 = t0, t1

    def __call__(self, x, out=None):
        x = _arrays.materialize(x)
        re, im, shape = _arrays.complex_view(x)
        if len(shape) < 2 or shape[-1] != self.t0.n or shape[-2] != self.t1.n:
            raise ValueError(f"array shape {shape} does not end in ({self.t1.n}, {self.t0.n})")
        prec = _arrays.precision_of_complex(x)
        if out is None:
            out = _arrays.empty_like_complex(x)
        _arrays.check_complex_out(out, shape, prec)
        ore, oim = _arrays.ptrs(out)
        batch = _arrays.batch_of(shape[:-1])
        _abi.check(_abi.lib().af_execute_2d(self.t0.plan(prec).handle, self.t1.plan(prec).handle, batch, re, im, ore,
                                            oim, _arrays.stream_handle()), "af_execute_2d")
        return out


class Transform3d:
    """``Preprocessed.transform3d (SubTransformTriple t0 t1 t2)`` (Preprocessed.hs:150-161):
    t0, t1, t2 along the last, second-to-last and third-to-last axes."""

    def __init__(self, t0: Transform, t1: Transform, t2: Transform):
        self.t0, self.t1, self.t2 = t0, t1, t2

    def __call__(self, x, out=None):
        x = _arrays.materialize(x)
        re, im, shape = _arrays.complex_view(x)
        if len(shape) < 3 or shape[-3:] != (self.t2.n, self.t1.n, self.t0.n):
            raise ValueError(f"array shape {shape} does not end in ({self.t2.n}, {self.t1.n}, {self.t0.n})")
        prec = _arrays.precision_of_complex(x)
        if out is None:
            out = _arrays.empty_like_complex(x)
        _arrays.check_complex_out(out, shape, prec)
        ore, oim = _arrays.ptrs(out)
        batch = _arrays.batch_of(shape[:-2])
        _abi.check(_abi.lib().af_execute_3d(self.t0.plan(prec).handle, self.t1.plan(prec).handle,
                                            self.t2.plan(prec).handle, batch, re, im, ore, oim,
                                            _arrays.stream_handle()), "af_execute_3d")
        return out


def transform2d(t0: Transform, t1: Transform) -> Transform2d:
    return Transform2d(t0, t1)


def transform3d(t0: Transform, t1: Transform, t2: Transform) -> Transform3d:
    return Transform3d(t0, t1, t2)


def _transpose_last2(x, rows, cols, out_shape):
    x = _arrays.materialize(x)
    re, im, shape = _arrays.complex_view(x)
    prec = _arrays.precision_of_complex(x)
    out = _arrays.empty_like_complex(x, out_shape)
    ore, oim = _arrays.ptrs(out)
    batch = 1
    for d in shape:
        batch *= d
    batch //= max(rows * cols, 1)
    _abi.check(_abi.lib().af_transpose(prec, batch, rows, cols, re, im, ore, oim, _arrays.stream_handle()),
               "af_transpose")
    return out


def transpose(x):
    """LinAlg.transpose of the last two axes: (..., h, w) -> (..., w, h)."""
    shape = tuple(x[0].shape if isinstance(x, tuple) else x.shape)
    return _transpose_last2(x, shape[-2], shape[-1], shape[:-2] + (shape[-1], shape[-2]))


def cycle_dim3(x):
    """Fourier.cycleDim3 (Private.hs:414-425): (..., k, m, n) -> (..., n, k, m)."""
    shape = tuple(x[0].shape if isinstance(x, tuple) else x.shape)
    k, m, n = shape[-3:]
    return _transpose_last2(x, k * m, n, shape[:-3] + (n, k, m))


def wisdom_export() -> str:
    """The autotuner's choices in this process, as text."""
    need = ctypes.c_size_t(0)
    _abi.check(_abi.lib().af_wisdom_export(None, 0, ctypes.byref(need)), "af_wisdom_export")
    buf = ctypes.create_string_buffer(max(need.value, 1))
    _abi.check(_abi.lib().af_wisdom_export(buf, len(buf), ctypes.byref(need)), "af_wisdom_export")
    return buf.value.decode()


def wisdom_import(text: str) -> None:
    """Adds exported autotuner choices to this process's wisdom."""
    _abi.check(_abi.lib().af_wisdom_import(text.encode()), "af_wisdom_import")


def transform(sign: Sign, n: int) -> Transform:
    """Planned.transform (Planned.hs:103-106)."""
    return Transform(sign, n)


def transform_decompose(sign: Sign, n: int) -> Transform:
    """Planned.transformDecompose (Planned.hs:114-119): no Bluestein."""
    return Transform(sign, n, _abi.ALG_DECOMPOSE)


def transform_chirp2(sign: Sign, n: int) -> Transform:
    """Planned.transformChirp2 (Planned.hs:704-708): Bluestein, power-of-two pad."""
    return Transform(sign, n, _abi.ALG_CHIRP2)


def transform_chirp235(sign: Sign, n: int) -> Transform:
    """Planned.transformChirp235 (Planned.hs:716-720): Bluestein, 5-smooth pad."""
    return Transform(sign, n, _abi.ALG_CHIRP235)


def ceiling_power_of_two(n: int) -> int:
    """NumberTheory.ceilingPowerOfTwo: the least power of two >= n."""
    return int(_abi.lib().af_ceiling_power_of_two(int(n)))


def ceiling_5_smooth(n: int) -> int:
    """NumberTheory.ceiling5Smooth: the least 5-smooth number >= n."""
    return int(_abi.lib().af_ceiling_5smooth(int(n)))


class Planner:
    """A planning rule ``Integer -> State PlanMap Plan`` of Planned.hs.

    The device planner knows the reference's three rules: ``planWithMapUpdate``
    (Rader, or Bluestein where Rader would nest; Planned.hs:215-240),
    ``planDecomposeWithMapUpdate`` (never Bluestein; 279-293) and
    ``planChirpWithMapUpdate padLength`` (Bluestein with a pad of
    ceilingPowerOfTwo or ceiling5Smooth of 2n-1; 731-737).  Each is one
    ``algorithm`` of the C ABI's plan descriptor.  Calling a planner gives the
    text of the device plan it makes for a length."""

    def __init__(self, name: str, algorithm: int):
        self.name = name
        self.algorithm = algorithm

    def __call__(self, n: int, precision=_abi.F32) -> str:
        return describe(int(n), precision=precision, algorithm=self.algorithm)

    def __repr__(self):
        return f"Planner({self.name})"


plan_with_map_update = Planner("planWithMapUpdate", _abi.ALG_AUTO)
plan_decompose_with_map_update = Planner("planDecomposeWithMapUpdate", _abi.ALG_DECOMPOSE)


def plan_chirp_with_map_update(pad_length) -> Planner:
    """Planned.planChirpWithMapUpdate (Planned.hs:731-737) for the two pad
    functions the reference's own entry points pass (transformChirp2 and
    transformChirp235, Planned.hs:704-720)."""
    if pad_length is ceiling_power_of_two:
        return Planner("planChirpWithMapUpdate ceilingPowerOfTwo", _abi.ALG_CHIRP2)
    if pad_length is ceiling_5_smooth:
        return Planner("planChirpWithMapUpdate ceiling5Smooth", _abi.ALG_CHIRP235)
    raise ValueError("the device planner pads Bluestein transforms with ceiling_power_of_two "
                     "or ceiling_5_smooth only")


def plan(n: int, precision=_abi.F32) -> str:
    """Planned.plan (Planned.hs:201-203): the default plan for a length, as the
    text of the device plan."""
    return plan_with_map_update(n, precision)


def transform_with_planner(planner: Planner, sign: Sign, n: int) -> Transform:
    """Planned.transformWithPlanner (Planned.hs:121-130).  A planner is one of
    the reference's planning rules above; an arbitrary function from lengths
    to plans has no device counterpart and is refused."""
    if not isinstance(planner, Planner):
        raise TypeError("planner must be plan_with_map_update, plan_decompose_with_map_update "
                        "or plan_chirp_with_map_update(pad)")
    return Transform(sign, n, planner.algorithm)


class Cache:
    """Planned.Cache (Planned.hs:300-309): everything a transform of one
    length and sign precomputes.  The device tables are built on first use for
    each precision and shared with every other transform of the same length,
    sign and algorithm in this process."""

    def __init__(self, sign: Sign, n: int, algorithm: int = _abi.ALG_AUTO, adjoint_of: "Cache" = None):
        self.sign = sign if isinstance(sign, Sign) else Sign(int(sign))
        self.n = int(n)
        self.algorithm = algorithm
        self.adjoint_of = adjoint_of

    def transform(self) -> Transform:
        if self.adjoint_of is None:
            return Transform(self.sign, self.n, self.algorithm)
        return _AdjointTransform(self.adjoint_of.transform())


class _AdjointTransform(Transform):
    """The transform of the opposite sign run as the adjoint of another
    transform's plan, so the two share every table: the unnormalised DFT
    matrix is symmetric, so its adjoint is its conjugate, the DFT of the
    opposite sign."""

    def __init__(self, of: Transform):
        super().__init__(of.sign.negate(), of.n, of.algorithm)
        self._of = of

    def __call__(self, x, out=None):
        return self._of.vjp(x, out)

    # plan() stays the opposite-sign plan of its own: 2-D and 3-D
    # compositions run plans directly and need the right sign.

    def tune(self, batch, precision="f32", layout="planar"):
        raise NotImplementedError("tune the forward transform of the pair")

    def vjp(self, g, out=None):
        return self._of(g, out)


def cache(sign: Sign, n: int) -> Cache:
    """Planned.cache (Planned.hs:316-329)."""
    return Cache(sign, n)


def cache_duplex(n: int):
    """Planned.cacheDuplex (Planned.hs:331-341): ``(cache inverse n, cache
    forward n)`` sharing their data.  The inverse runs as the adjoint of the
    forward transform's device plan, so one set of tables serves both."""
    fwd = Cache(Sign(-1), n)
    return Cache(Sign(1), n, adjoint_of=fwd), fwd


def transform_with_cache(c: Cache) -> Transform:
    """Planned.transformWithCache (Planned.hs:137-167)."""
    return c.transform()


def convolve_cyclic(n: int):
    """Planned.convolveCyclic (Planned.hs:779-792): normalised cyclic convolution.

    Signals must have equal size and must not be empty."""
    if n < 1:
        raise ValueError("convolveCyclic needs a non-empty length")

    def conv(x, y, out=None):
        x, y = _arrays.materialize(x), _arrays.materialize(y)
        xr, xi, xs = _arrays.complex_view(x, "x")
        yr, yi, ys = _arrays.complex_view(y, "y")
        if xs != ys or xs[-1] != n:
            raise ValueError("operands must have equal shape with inner extent n")
        prec = _arrays.precision_of_complex(x)
        if _arrays.precision_of_complex(y) != prec:
            raise TypeError("operands must have the same precision")
        if out is None:
            out = _arrays.empty_like_complex(x)
        _arrays.check_complex_out(out, xs, prec)
        ore, oim = _arrays.ptrs(out)
        p = get_plan(n, -1, _abi.CONV_CYCLIC, prec)
        _abi.check(_abi.lib().af_convolve_cyclic(p.handle, _arrays.batch_of(xs), xr, xi, yr, yi, ore, oim, None,
                                                 _arrays.stream_handle()), "af_convolve_cyclic")
        return out

    return conv
