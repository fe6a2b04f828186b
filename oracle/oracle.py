"""ctypes front-end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only the test-suite, ``__graft_entry__.smoke()`` (as the checker) and
``bench.py``'s ``cpu_baseline`` leg import this module.  The HIP product path
never does.  The oracle restates the reference's algorithms in C
(``fourier_oracle.c`` / ``fourier_oracle_body.h``; each C function cites the
reference file:line it follows).  Arrays are numpy arrays; complex data is
passed as Accelerate does, as planar real/imaginary component arrays.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_lib = None

AUTO, DECOMPOSE, CHIRP2, CHIRP235 = 0, 1, 2, 3

_i64 = ctypes.c_int64
_int = ctypes.c_int
_p = ctypes.c_void_p


def build() -> None:
    subprocess.run(["sh", os.path.join(_HERE, "build_oracle.sh")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_fermat_factors": (_i64, [_i64, _p, _i64]),
            "orc_multiplicative_generator": (_i64, [_i64]),
            "orc_ceiling_power_of_two": (_i64, [_i64]),
            "orc_ceiling_5smooth": (_i64, [_i64]),
            "orc_perm_multiplicative": (None, [_i64, _p]),
            "orc_perm_inverse": (None, [_i64, _p, _p]),
            "orc_perm_reverse": (None, [_i64, _p, _p]),
            "orc_skew_grid": (None, [_i64, _i64, _p]),
            "orc_skew_grid_inv": (None, [_i64, _i64, _p]),
            "orc_plan_string": (_int, [_i64, _int, ctypes.c_char_p, _i64]),
            "orc_plan_count_primes": (_int, [_i64]),
            "orc_transform_f64": (_int, [_int, _int, _i64, _i64, _p, _p, _p, _p]),
            "orc_transform_f32": (_int, [_int, _int, _i64, _i64, _p, _p, _p, _p]),
            "orc_dft_bruteforce": (None, [_int, _i64, _i64, _p, _p, _p, _p]),
            "orc_untangle_spectra_f64": (_int, [_i64, _i64, _p, _p, _p, _p, _p, _p]),
            "orc_entangle_spectra_f64": (_int, [_i64, _i64, _p, _p, _p, _p, _p, _p]),
            "orc_untangle_spectra_f32": (_int, [_i64, _i64, _p, _p, _p, _p, _p, _p]),
            "orc_entangle_spectra_f32": (_int, [_i64, _i64, _p, _p, _p, _p, _p, _p]),
            "orc_to_spectrum_f64": (_int, [_int, _i64, _i64, _p, _p, _p]),
            "orc_from_spectrum_f64": (_int, [_int, _i64, _i64, _p, _p, _p]),
            "orc_to_spectrum_f32": (_int, [_int, _i64, _i64, _p, _p, _p]),
            "orc_from_spectrum_f32": (_int, [_int, _i64, _i64, _p, _p, _p]),
            "orc_convolve_cyclic_f64": (_int, [_i64, _i64, _p, _p, _p, _p, _p, _p]),
            "orc_karatsuba_i64": (None, [_i64, _i64, _p, _p, _p]),
            "orc_karatsuba_rec_i64": (None, [_i64, _i64, _p, _p, _p]),
            "orc_karatsuba_f64": (None, [_i64, _i64, _p, _p, _p]),
            "orc_schoolbook_i64": (None, [_i64, _i64, _p, _p, _p]),
            "orc_karatsuba_ragged_i64": (None, [_i64, _p, _p, _p, _p]),
            "orc_cyclic_karatsuba_i64": (None, [_i64, _i64, _p, _p, _p]),
            "orc_complex_karatsuba_f64": (None, [_int, _i64, _i64, _p, _p, _p, _p, _p, _p]),
            "orc_complex_karatsuba_i64": (None, [_int, _i64, _i64, _p, _p, _p, _p, _p, _p]),
            "orc_num_threads": (_int, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


# ---- number theory / permutations / index maps ----

def fermat_factors(n: int) -> list[tuple[int, int]]:
    cnt = lib().orc_fermat_factors(n, None, 0)
    buf = np.zeros(2 * max(cnt, 1), dtype=np.int64)
    lib().orc_fermat_factors(n, _ptr(buf), cnt)
    return [(int(buf[2 * i]), int(buf[2 * i + 1])) for i in range(cnt)]


def multiplicative_generator(p: int) -> int:
    return int(lib().orc_multiplicative_generator(p))


def ceiling_power_of_two(n: int) -> int:
    return int(lib().orc_ceiling_power_of_two(n))


def ceiling_5smooth(n: int) -> int:
    return int(lib().orc_ceiling_5smooth(n))


def perm_multiplicative(p: int) -> np.ndarray:
    out = np.zeros(p - 1, dtype=np.int64)
    lib().orc_perm_multiplicative(p, _ptr(out))
    return out


def perm_inverse(perm) -> np.ndarray:
    perm = _c(perm, np.int64)
    out = np.zeros_like(perm)
    lib().orc_perm_inverse(len(perm), _ptr(perm), _ptr(out))
    return out


def perm_reverse(perm) -> np.ndarray:
    perm = _c(perm, np.int64)
    out = np.zeros_like(perm)
    lib().orc_perm_reverse(len(perm), _ptr(perm), _ptr(out))
    return out


def skew_grid(m: int, n: int) -> np.ndarray:
    out = np.zeros(m * n, dtype=np.int64)
    lib().orc_skew_grid(m, n, _ptr(out))
    return out


def skew_grid_inv(n: int, m: int) -> np.ndarray:
    out = np.zeros(m * n, dtype=np.int64)
    lib().orc_skew_grid_inv(n, m, _ptr(out))
    return out


def plan_string(n: int, algorithm: int = AUTO) -> str:
    buf = ctypes.create_string_buffer(1 << 16)
    if lib().orc_plan_string(n, algorithm, buf, len(buf)) != 0:
        raise ValueError("plan string too long")
    return buf.value.decode()


def plan_count_primes(n: int) -> int:
    return int(lib().orc_plan_count_primes(n))


# ---- transforms ----

def transform(sign: int, x: np.ndarray, algorithm: int = AUTO, precision: str = "f64") -> np.ndarray:
    """Planned.transform sign n over the last axis of a complex array."""
    x = np.asarray(x)
    n = x.shape[-1] if x.ndim else 0
    rows = int(np.prod(x.shape[:-1])) if x.ndim > 1 else 1
    rt = np.float64 if precision == "f64" else np.float32
    re = _c(x.real, rt)
    im = _c(x.imag, rt)
    ore = np.empty_like(re)
    oim = np.empty_like(im)
    fn = lib().orc_transform_f64 if precision == "f64" else lib().orc_transform_f32
    if fn(algorithm, sign, n, rows, _ptr(re), _ptr(im), _ptr(ore), _ptr(oim)) != 0:
        raise ValueError("invalid arguments")
    ct = np.complex128 if precision == "f64" else np.complex64
    out = np.empty(x.shape, dtype=ct)
    out.real = ore
    out.imag = oim
    return out


def dft_bruteforce(sign: int, x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.complex128)
    n = x.shape[-1]
    rows = int(np.prod(x.shape[:-1])) if x.ndim > 1 else 1
    re = _c(x.real, np.float64)
    im = _c(x.imag, np.float64)
    ore = np.empty_like(re)
    oim = np.empty_like(im)
    lib().orc_dft_bruteforce(sign, n, rows, _ptr(re), _ptr(im), _ptr(ore), _ptr(oim))
    return ore + 1j * oim


def to_spectrum(x: np.ndarray, sub_sign: int = -1, precision: str = "f64") -> np.ndarray:
    rt = np.float64 if precision == "f64" else np.float32
    x = _c(x, rt)
    n = x.shape[-1]
    rows = int(np.prod(x.shape[:-1])) if x.ndim > 1 else 1
    ore = np.empty_like(x)
    oim = np.empty_like(x)
    fn = lib().orc_to_spectrum_f64 if precision == "f64" else lib().orc_to_spectrum_f32
    if fn(sub_sign, n, rows, _ptr(x), _ptr(ore), _ptr(oim)) != 0:
        raise ValueError("invalid arguments")
    return ore + 1j * oim


def untangle_spectra(z: np.ndarray, precision: str = "f64"):
    """Real.untangleSpectra (Real.hs:168-187) over the last axis: (a, b)."""
    rt = np.float64 if precision == "f64" else np.float32
    zre, zim = _c(np.real(z), rt), _c(np.imag(z), rt)
    n = zre.shape[-1]
    rows = int(np.prod(zre.shape[:-1])) if zre.ndim > 1 else 1
    out = [np.empty_like(zre) for _ in range(4)]
    fn = lib().orc_untangle_spectra_f64 if precision == "f64" else lib().orc_untangle_spectra_f32
    if fn(n, rows, _ptr(zre), _ptr(zim), *[_ptr(o) for o in out]) != 0:
        raise ValueError("invalid arguments")
    return out[0] + 1j * out[1], out[2] + 1j * out[3]


def entangle_spectra(c: np.ndarray, d: np.ndarray, precision: str = "f64") -> np.ndarray:
    """Real.entangleSpectra (Real.hs:208-236): c + d i."""
    rt = np.float64 if precision == "f64" else np.float32
    parts = [_c(np.real(c), rt), _c(np.imag(c), rt), _c(np.real(d), rt), _c(np.imag(d), rt)]
    n = parts[0].shape[-1]
    rows = int(np.prod(parts[0].shape[:-1])) if parts[0].ndim > 1 else 1
    zre, zim = np.empty_like(parts[0]), np.empty_like(parts[0])
    fn = lib().orc_entangle_spectra_f64 if precision == "f64" else lib().orc_entangle_spectra_f32
    if fn(n, rows, *[_ptr(q) for q in parts], _ptr(zre), _ptr(zim)) != 0:
        raise ValueError("invalid arguments")
    return zre + 1j * zim


def two_to_spectrum(sign: int, x: np.ndarray, y: np.ndarray, precision: str = "f64"):
    """Real.twoToSpectrum (Planned.transform sign n) (Real.hs:129-135): the
    spectra of two real signals from one complex transform of x + i y."""
    return untangle_spectra(transform(sign, np.asarray(x) + 1j * np.asarray(y), precision=precision), precision)


def two_from_spectrum(sign: int, a: np.ndarray, b: np.ndarray, precision: str = "f64"):
    """Real.twoFromSpectrum (Planned.transform sign n) (Real.hs:190-196): the
    real and imaginary parts of the transform of the entangled spectra."""
    z = transform(sign, entangle_spectra(a, b, precision), precision=precision)
    return np.real(z), np.imag(z)


def from_spectrum(s: np.ndarray, sub_sign: int = 1, precision: str = "f64") -> np.ndarray:
    rt = np.float64 if precision == "f64" else np.float32
    re = _c(np.real(s), rt)
    im = _c(np.imag(s), rt)
    n = re.shape[-1]
    rows = int(np.prod(re.shape[:-1])) if re.ndim > 1 else 1
    y = np.empty_like(re)
    fn = lib().orc_from_spectrum_f64 if precision == "f64" else lib().orc_from_spectrum_f32
    if fn(sub_sign, n, rows, _ptr(re), _ptr(im), _ptr(y)) != 0:
        raise ValueError("invalid arguments")
    return y


def convolve_cyclic(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.complex128)
    y = np.asarray(y, dtype=np.complex128)
    n = x.shape[-1]
    rows = int(np.prod(x.shape[:-1])) if x.ndim > 1 else 1
    a, b, c, d = (_c(v, np.float64) for v in (x.real, x.imag, y.real, y.imag))
    ore = np.empty_like(a)
    oim = np.empty_like(a)
    if lib().orc_convolve_cyclic_f64(n, rows, _ptr(a), _ptr(b), _ptr(c), _ptr(d), _ptr(ore), _ptr(oim)) != 0:
        raise ValueError("invalid arguments")
    return ore + 1j * oim


# ---- Karatsuba ----

def _kara(fn, x, y, dtype, outlen=None):
    x = _c(x, dtype)
    y = _c(y, dtype)
    n = x.shape[-1]
    rows = int(np.prod(x.shape[:-1])) if x.ndim > 1 else 1
    zl = (2 * n - 1 if n > 0 else 0) if outlen is None else outlen
    z = np.zeros(x.shape[:-1] + (zl,), dtype=dtype)
    fn(n, rows, _ptr(x), _ptr(y), _ptr(z))
    return z


def karatsuba(x, y):
    return _kara(lib().orc_karatsuba_i64, x, y, np.int64)


def karatsuba_rec(x, y):
    return _kara(lib().orc_karatsuba_rec_i64, x, y, np.int64)


def karatsuba_f64(x, y):
    return _kara(lib().orc_karatsuba_f64, x, y, np.float64)


def schoolbook(x, y):
    return _kara(lib().orc_schoolbook_i64, x, y, np.int64)


def cyclic_karatsuba(x, y):
    x = np.asarray(x)
    return _kara(lib().orc_cyclic_karatsuba_i64, x, y, np.int64, outlen=x.shape[-1])


def karatsuba_ragged(offsets, x, y):
    offsets = _c(offsets, np.int64)
    x = _c(x, np.int64)
    y = _c(y, np.int64)
    npairs = len(offsets) - 1
    z = np.zeros(2 * int(offsets[-1]) - npairs, dtype=np.int64)
    lib().orc_karatsuba_ragged_i64(npairs, _ptr(offsets), _ptr(x), _ptr(y), _ptr(z))
    return z


def complex_karatsuba(x, y, cyclic: bool = False):
    x = np.asarray(x, dtype=np.complex128)
    y = np.asarray(y, dtype=np.complex128)
    n = x.shape[-1]
    rows = int(np.prod(x.shape[:-1])) if x.ndim > 1 else 1
    ol = n if cyclic else max(2 * n - 1, 0)
    a, b, c, d = (_c(v, np.float64) for v in (x.real, x.imag, y.real, y.imag))
    ore = np.zeros(x.shape[:-1] + (ol,))
    oim = np.zeros(x.shape[:-1] + (ol,))
    lib().orc_complex_karatsuba_f64(1 if cyclic else 0, n, rows, _ptr(a), _ptr(b), _ptr(c), _ptr(d),
                                    _ptr(ore), _ptr(oim))
    return ore + 1j * oim


def complex_karatsuba_i64(xr, xi, yr, yi, cyclic: bool = False):
    """Convolution.complex karatsuba (Adhoc.hs:114-133) on Int64 (re, im) parts:
    returns (re, im), wrap-around arithmetic."""
    xr, xi, yr, yi = (_c(v, np.int64) for v in (xr, xi, yr, yi))
    n = xr.shape[-1]
    rows = int(np.prod(xr.shape[:-1])) if xr.ndim > 1 else 1
    ol = n if cyclic else max(2 * n - 1, 0)
    ore = np.zeros(xr.shape[:-1] + (ol,), dtype=np.int64)
    oim = np.zeros(xr.shape[:-1] + (ol,), dtype=np.int64)
    lib().orc_complex_karatsuba_i64(1 if cyclic else 0, n, rows, _ptr(xr), _ptr(xi), _ptr(yr), _ptr(yi),
                                    _ptr(ore), _ptr(oim))
    return ore, oim


def num_threads() -> int:
    return int(lib().orc_num_threads())


# ---- multi-dimensional transforms: the reference's compositions of 1-D ones ----

def transform2d(sign0: int, sign1: int, x: np.ndarray, precision: str = "f64") -> np.ndarray:
    """Preprocessed.transform2d (SubTransformPair (transform sign0 w) (transform sign1 h))
    (Preprocessed.hs:141-148): transpose . t1 . transpose . t0 over the last two axes."""
    y = transform(sign0, np.asarray(x), precision=precision)
    y = transform(sign1, np.ascontiguousarray(np.swapaxes(y, -1, -2)), precision=precision)
    return np.ascontiguousarray(np.swapaxes(y, -1, -2))


def cycle_dim3(x: np.ndarray) -> np.ndarray:
    """Fourier.cycleDim3 (Private.hs:414-425): (..., k, m, n) -> (..., n, k, m)."""
    return np.ascontiguousarray(np.moveaxis(np.asarray(x), -1, -3))


def transform3d(sign0: int, sign1: int, sign2: int, x: np.ndarray, precision: str = "f64") -> np.ndarray:
    """Preprocessed.transform3d (SubTransformTriple t0 t1 t2) (Preprocessed.hs:150-161):
    cycleDim3 . t2 . cycleDim3 . t1 . cycleDim3 . t0."""
    y = transform(sign0, np.asarray(x), precision=precision)
    y = transform(sign1, cycle_dim3(y), precision=precision)
    y = transform(sign2, cycle_dim3(y), precision=precision)
    return cycle_dim3(y)


def reverse2d(x: np.ndarray) -> np.ndarray:
    """Cyclic.reverse2d (Cyclic.hs:30-39): (y, x) -> ((-y) mod h, (-x) mod w)."""
    x = np.asarray(x)
    h, w = x.shape[-2:]
    return x[..., (-np.arange(h)) % h, :][..., (-np.arange(w)) % w]


def untangle_spectra_2d(z: np.ndarray):
    """Real.untangleSpectra2d (Real.hs:175-180) with untangleCoefficient (182-187)."""
    bc = np.conj(reverse2d(z))
    return (z + bc) / 2, (z - bc) * (-0.5j)


def two_to_spectrum_2d(sign0: int, sign1: int, x: np.ndarray, y: np.ndarray):
    """Real.twoToSpectrum2d (Real.hs:138-145)."""
    return untangle_spectra_2d(transform2d(sign0, sign1, np.asarray(x) + 1j * np.asarray(y)))


def two_from_spectrum_2d(sign0: int, sign1: int, a: np.ndarray, b: np.ndarray):
    """Real.twoFromSpectrum2d (Real.hs:199-206) with entangleSpectra2d (208-216)."""
    z = transform2d(sign0, sign1, entangle_spectra(a, b))
    return np.real(z), np.imag(z)
