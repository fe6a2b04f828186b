"""Pins the oracle (the C restatement of the reference's algorithms) before it
is trusted as the GPU path's checker.

Every oracle function is checked against something independent of it: mpmath
DFTs at 40 digits (tests/golden), a long-double brute-force DFT with exact
index reduction, numpy.fft (pocketfft), closed forms (basis vectors), sympy's
primitive roots, brute-force searches, exact Python-integer convolutions, and
the algebraic properties the reference's own test-suite states
(test/Test/Data/Array/Accelerate/Fourier.hs:298-587).
"""
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dft_mpmath.npz")
ALGS = [O.AUTO, O.DECOMPOSE, O.CHIRP2, O.CHIRP235]


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / (nb if nb else 1.0)


def rng_c(seed, shape):
    r = np.random.default_rng(seed)
    return r.uniform(-1, 1, shape) + 1j * r.uniform(-1, 1, shape)


# ---------------------------------------------------------------------------
# golden vectors (mpmath) and brute force
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


@pytest.mark.parametrize("alg", ALGS)
def test_golden_mpmath(golden, alg):
    for n in golden["lengths"]:
        x = golden[f"x{n}"]
        for sign, key in ((-1, "fwd"), (1, "inv")):
            got = O.transform(sign, x, algorithm=alg)
            assert rel(got, golden[f"{key}{n}"]) < 1e-13, (n, sign, alg, O.plan_string(int(n), alg))


@pytest.mark.parametrize("alg", ALGS)
def test_bruteforce_small(alg):
    for n in range(0, 90):
        x = rng_c(n, (2, n))
        for sign in (-1, 1):
            got = O.transform(sign, x, algorithm=alg)
            if n == 0:
                assert got.shape == (2, 0)
                continue
            assert rel(got, O.dft_bruteforce(sign, x)) < 1e-13, (n, sign, alg)


def tol_f64(n, alg=O.AUTO):
    """The reference's Bluestein chirp phase pi k^2 / n is not reduced modulo
    2n (Private.hs:428-436), so its FP64 error grows like eps * n."""
    return 1e-13 + (4e-15 * n if O.plan_string(n, alg).startswith("chirp") else 0.0)


@pytest.mark.parametrize("n", [256, 1000, 1024, 2310, 4096, 4099, 10007, 65521, 65536])
def test_numpy_large(n):
    x = rng_c(n, (2, n))
    assert rel(O.transform(-1, x), np.fft.fft(x)) < tol_f64(n)
    assert rel(O.transform(1, x), np.fft.ifft(x) * n) < tol_f64(n)


def test_bruteforce_is_the_dft():
    """The brute-force checker itself against numpy on a few lengths."""
    for n in (1, 5, 12, 31):
        x = rng_c(100 + n, (n,))
        assert rel(O.dft_bruteforce(-1, x), np.fft.fft(x)) < 1e-14


def test_basis_vector():
    """Fourier.hs:367-376: the inverse transform of impulse k is cis(2 pi k j / n)."""
    for n in (1, 2, 7, 12, 47, 97, 100):
        for k in sorted({0, 1 % n, n // 2, n - 1}):
            e = np.zeros(n, complex)
            e[k] = 1
            j = np.arange(n)
            expect = np.exp(2j * np.pi * ((k * j) % n) / n)
            assert np.max(np.abs(O.transform(1, e) - expect)) < 1e-12


def test_precision_f32_restatement():
    for n in (64, 97, 1000, 4096):
        x = rng_c(n, (2, n))
        assert rel(O.transform(-1, x, precision="f32"), np.fft.fft(x)) < 5e-6


# ---------------------------------------------------------------------------
# the reference's algebraic properties (Fourier.hs:355-417)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 3, 6, 9, 15, 31, 47, 64, 97])
def test_properties(n):
    x = rng_c(n, (n,))
    y = rng_c(n + 1, (n,))
    a = 0.3 - 0.7j
    for s in (-1, 1):
        T = lambda v: O.transform(s, v)
        assert np.allclose(T(a * x), a * T(x), atol=1e-10)                      # homogeneity
        assert np.allclose(T(x) + T(y), T(x + y), atol=1e-10)                   # additivity
        rev = lambda v: np.roll(v[::-1], 1)                                     # Cyclic.reverse
        assert np.allclose(rev(T(T(x))) / n, x, atol=1e-10)                     # double = reverse
        assert np.allclose(rev(T(x)), T(rev(x)), atol=1e-10)                    # transform of reverse
        assert np.allclose(np.conj(T(x)), T(np.conj(rev(x))), atol=1e-10)       # conjugate
        assert abs(np.linalg.norm(T(x)) - np.sqrt(n) * np.linalg.norm(x)) < 1e-10  # isometry
        assert abs(np.vdot(T(y), T(x)) - n * np.vdot(y, x)) < 1e-9             # unitarity
    assert np.allclose(O.transform(-1, O.transform(1, x)) / n, x, atol=1e-10)   # inverse


# ---------------------------------------------------------------------------
# planner and number theory
# ---------------------------------------------------------------------------

def test_plan_strings_hand_checked():
    """Plan trees derived by hand from Planned.hs:201-293 and NumberTheory.hs."""
    assert O.plan_string(1024) == "sr1024(sr512(sr256(sr128(sr64(sr32(sr16(sr8(c4))))))))"
    assert O.plan_string(9) == "ct9(c3,c3)"
    assert O.plan_string(15) == "gt15(c3,c5)"
    assert O.plan_string(45) == "gt45(c5,ct9(c3,c3))"
    assert O.plan_string(7) == "rader7(sr6(c3))"
    assert O.plan_string(47) == "chirp47(sr96(sr48(sr24(sr12(sr6(c3))))))"  # Rader depth 3 -> Bluestein
    assert O.plan_string(100003).startswith("chirp100003(sr202500(")
    assert O.plan_string(0) == "id0" and O.plan_string(1) == "id1"
    assert O.plan_string(5) == "c5"
    assert O.plan_string(47, O.DECOMPOSE).startswith("rader47(sr46(rader23(")
    assert O.plan_string(10, O.CHIRP2) == "chirp10(sr32(sr16(sr8(c4))))"
    assert O.plan_string(10, O.CHIRP235) == "chirp10(sr20(sr10(c5)))"


def test_count_primes():
    assert O.plan_count_primes(7) == 1
    assert O.plan_count_primes(47) == 3
    assert O.plan_count_primes(1024) == 0


def test_fermat_factors_bruteforce():
    for n in range(3, 400, 2):
        pairs = O.fermat_factors(n)
        divs = [(d, n // d) for d in range(1, int(np.sqrt(n)) + 1) if n % d == 0]
        assert sorted(pairs) == sorted(divs), n
        assert pairs[0] == max(divs)  # the pair closest to sqrt n comes first


def test_multiplicative_generator_sympy():
    from sympy import primerange
    from sympy.ntheory import primitive_root
    for p in list(primerange(3, 600)) + [65521, 100003]:
        assert O.multiplicative_generator(p) == primitive_root(p), p


def test_ceiling_smooth_bruteforce():
    def smooth5(m):
        for q in (2, 3, 5):
            while m % q == 0:
                m //= q
        return m == 1
    for n in list(range(1, 700)) + [2 * 100003 - 1, 2 * 1000003 - 1]:
        m = n
        while not smooth5(m):
            m += 1
        assert O.ceiling_5smooth(n) == m, n
        assert O.ceiling_power_of_two(n) == 1 << (n - 1).bit_length()


def test_permutations():
    for p in (3, 5, 7, 13, 97, 257):
        perm = O.perm_multiplicative(p)
        g = O.multiplicative_generator(p)
        assert np.array_equal(perm, [pow(g, i, p) - 1 for i in range(p - 1)])
        inv = O.perm_inverse(perm)
        assert np.array_equal(inv[perm], np.arange(p - 1))
        rev = O.perm_reverse(perm)
        assert np.array_equal(rev, perm[(-np.arange(p - 1)) % (p - 1)])
    # Planned.hs:632-643: generator 3 of Z7, permutation 1 3 2 6 4 5
    assert list(O.perm_multiplicative(7) + 1) == [1, 3, 2, 6, 4, 5]


def test_good_thomas_maps_are_crt_bijections():
    for n, m in ((3, 5), (5, 9), (7, 9), (4, 15)):
        g = O.skew_grid(m, n)
        assert sorted(g) == list(range(n * m))
        inv = O.skew_grid_inv(n, m)
        assert sorted(inv) == list(range(n * m))
        for q in range(n * m):
            a, b = divmod(int(inv[q]), m)
            assert a == q % n and b == q % m


# ---------------------------------------------------------------------------
# real transforms, convolutions
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n", [2, 4, 6, 10, 64, 100, 1024])
def test_real_spectrum(n):
    x = np.random.default_rng(n).standard_normal((3, n))
    assert rel(O.to_spectrum(x), np.fft.fft(x)) < 1e-13                         # Fourier.hs:463-469
    S = np.fft.fft(x)
    assert rel(O.from_spectrum(S), n * x) < 1e-13                               # Fourier.hs:470-477
    assert rel(O.from_spectrum(O.to_spectrum(x)) / n, x) < 1e-13                # Fourier.hs:478-486


@pytest.mark.parametrize("n", [2, 4, 6, 10, 64, 100, 1024])
def test_real_packing_f32(n):
    """The FP32 instantiations of toSpectrum / fromSpectrum (Real.hs:49-107) and
    untangleSpectra / entangleSpectra (Real.hs:168-236) against the same closed
    forms as the FP64 ones, at FP32 tolerances: numpy's FP64 spectrum, n x for
    the inverse, (z + conj z_-k) / 2 and (z - conj z_-k)(-i/2) on arbitrary
    input, and a + i b for the entangled pair."""
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal((3, n)).astype(np.float32)
    S = np.fft.fft(x.astype(np.float64))
    assert rel(O.to_spectrum(x, precision="f32"), S) < 2e-6
    assert rel(O.from_spectrum(S, precision="f32"), n * x) < 2e-6
    z = (rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))).astype(np.complex64)
    ua, ub = O.untangle_spectra(z, precision="f32")
    zr = np.conj(z[:, (-np.arange(n)) % n]).astype(np.complex128)
    assert rel(ua, (z + zr) / 2) < 1e-6
    assert rel(ub, (z - zr) * (-0.5j)) < 1e-6
    a, b = S[:2], np.fft.fft(rng.standard_normal((2, n)))
    assert rel(O.entangle_spectra(a, b, precision="f32"), a + 1j * b) < 1e-6


def test_convolve_cyclic_numpy():
    for n in (1, 2, 7, 16, 47, 100):
        x, y = rng_c(n, (2, n)), rng_c(n + 9, (2, n))
        ref = np.fft.ifft(np.fft.fft(x) * np.fft.fft(y))
        assert rel(O.convolve_cyclic(x, y), ref) < 1e-13


def test_karatsuba_golden(golden):
    for L in (1, 37, 200):
        a, b = golden[f"ka{L}"][None], golden[f"kb{L}"][None]
        assert np.array_equal(O.karatsuba(a, b)[0], golden[f"kz{L}"])
        assert np.array_equal(O.karatsuba_rec(a, b)[0], golden[f"kz{L}"])
        assert np.array_equal(O.schoolbook(a, b)[0], golden[f"kz{L}"])


def test_karatsuba_rec_vs_loop():
    """Fourier.hs:436-442, on Int64 and on every length up to 70."""
    r = np.random.default_rng(5)
    for L in range(0, 71):
        x = r.integers(-2 ** 40, 2 ** 40, (2, L))
        y = r.integers(-2 ** 40, 2 ** 40, (2, L))
        k = O.karatsuba(x, y)
        assert k.shape == (2, max(2 * L - 1, 0))
        assert np.array_equal(k, O.karatsuba_rec(x, y))
        assert np.array_equal(k, O.schoolbook(x, y))


def test_karatsuba_cyclic_and_ragged():
    r = np.random.default_rng(6)
    L = 33
    x, y = r.integers(-99, 99, (3, L)), r.integers(-99, 99, (3, L))
    lin = O.schoolbook(x, y)
    cyc = lin[:, :L].copy()
    cyc[:, :L - 1] += lin[:, L:]
    assert np.array_equal(O.cyclic_karatsuba(x, y), cyc)
    lens = [1, 5, 64, 3]
    off = np.concatenate([[0], np.cumsum(lens)])
    xa, ya = r.integers(-99, 99, off[-1]), r.integers(-99, 99, off[-1])
    z = O.karatsuba_ragged(off, xa, ya)
    for p, L in enumerate(lens):
        o = 2 * off[p] - p
        assert np.array_equal(z[o:o + 2 * L - 1], np.convolve(xa[off[p]:off[p + 1]], ya[off[p]:off[p + 1]]))


def test_complex_karatsuba_matches_fft_convolution():
    """Fourier.hs:447-462: Karatsuba (linear and cyclic) against padded FFT convolution."""
    for n in (1, 4, 9, 30):
        x, y = rng_c(n, (n,)), rng_c(n + 1, (n,))
        assert rel(O.complex_karatsuba(x, y), np.convolve(x, y)) < 1e-13
        ref = np.fft.ifft(np.fft.fft(x) * np.fft.fft(y))
        assert rel(O.complex_karatsuba(x, y, cyclic=True), ref) < 1e-13


def test_planted_mistake_is_caught():
    """A plausible slip (the conjugate twiddle in a radix-2 level) fails the
    brute-force comparison; shown here on a modified restatement in numpy."""
    n = 16
    x = rng_c(1, (n,))
    e, o = np.fft.fft(x[0::2]), np.fft.fft(x[1::2])
    tw_bad = np.exp(+2j * np.pi * np.arange(n // 2) / n)
    bad = np.concatenate([e + tw_bad * o, e - tw_bad * o])
    assert rel(bad, O.dft_bruteforce(-1, x)) > 1e-3


@pytest.mark.parametrize("n", [1, 2, 7, 10, 97, 100, 1000])
def test_two_for_one_packing(n):
    """Real.hs:129-236 against closed forms and the reference's own properties
    (Fourier.hs:503-536): the untangled spectra of x + i y are the spectra of x
    and y (numpy and the oracle's transform), untangling an entangled pair of
    real spectra gives it back, and twoFromSpectrum with the inverse transform
    undoes twoToSpectrum up to the factor n."""
    rng = np.random.default_rng(n)
    x, y = rng.standard_normal((3, n)), rng.standard_normal((3, n))
    a, b = O.two_to_spectrum(-1, x, y)
    np.testing.assert_allclose(a, np.fft.fft(x, axis=1), rtol=0, atol=1e-11 * max(n, 1))
    np.testing.assert_allclose(b, np.fft.fft(y, axis=1), rtol=0, atol=1e-11 * max(n, 1))
    np.testing.assert_allclose(a, O.transform(-1, x + 0j), rtol=0, atol=1e-12 * max(n, 1))
    a2, b2 = O.untangle_spectra(O.entangle_spectra(a, b))
    np.testing.assert_allclose(a2, a, rtol=0, atol=1e-12 * max(n, 1))
    np.testing.assert_allclose(b2, b, rtol=0, atol=1e-12 * max(n, 1))
    xs, ys = O.two_from_spectrum(1, a, b)
    np.testing.assert_allclose(xs, n * x, rtol=0, atol=1e-10 * max(n, 1))
    np.testing.assert_allclose(ys, n * y, rtol=0, atol=1e-10 * max(n, 1))
    # the coefficient formulas themselves, on arbitrary (non-Hermitian) input:
    # (z + conj z_-k) / 2 and (z - conj z_-k) (-i / 2)
    z = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    ua, ub = O.untangle_spectra(z)
    zr = np.conj(z[:, (-np.arange(n)) % n])
    np.testing.assert_allclose(ua, (z + zr) / 2, rtol=0, atol=1e-15)
    np.testing.assert_allclose(ub, (z - zr) * (-0.5j), rtol=0, atol=1e-15)
    np.testing.assert_allclose(O.entangle_spectra(a, b), a + 1j * b, rtol=0, atol=1e-12 * max(n, 1))


@pytest.mark.parametrize("h,w", [(1, 1), (3, 5), (8, 6), (7, 16)])
def test_multidimensional(h, w):
    """Preprocessed.transform2d/3d, cycleDim3, reverse2d and the 2-D two-for-one
    packing (restated as the reference's compositions of 1-D transforms) against
    numpy's fft2/fftn, and the reference's transposition properties
    (Fourier.hs:537-586)."""
    rng = np.random.default_rng(h * 31 + w)
    x = rng.standard_normal((2, h, w)) + 1j * rng.standard_normal((2, h, w))
    np.testing.assert_allclose(O.transform2d(-1, -1, x), np.fft.fft2(x), rtol=0, atol=1e-12 * h * w)
    np.testing.assert_allclose(O.transform2d(1, 1, x), np.fft.ifft2(x) * h * w, rtol=0, atol=1e-12 * h * w)
    # transpose (transform2d (tw, th) x) == transform2d (th, tw) (transpose x)
    np.testing.assert_allclose(np.swapaxes(O.transform2d(-1, 1, x), -1, -2),
                               O.transform2d(1, -1, np.swapaxes(x, -1, -2)), rtol=0, atol=1e-12 * h * w)
    d = 3
    x3 = rng.standard_normal((2, d, h, w)) + 1j * rng.standard_normal((2, d, h, w))
    np.testing.assert_allclose(O.transform3d(-1, -1, -1, x3), np.fft.fftn(x3, axes=(-3, -2, -1)), rtol=0,
                               atol=1e-12 * d * h * w)
    np.testing.assert_allclose(O.cycle_dim3(O.transform3d(-1, 1, -1, x3)),
                               O.transform3d(1, -1, -1, O.cycle_dim3(x3)), rtol=0, atol=1e-12 * d * h * w)
    assert O.cycle_dim3(x3).shape == (2, w, d, h)
    assert np.array_equal(O.cycle_dim3(O.cycle_dim3(O.cycle_dim3(x3))), x3)
    # reverse2d: the spectrum of conj(signal) is conj(reverse2d(spectrum))
    np.testing.assert_allclose(np.fft.fft2(np.conj(x)), np.conj(O.reverse2d(np.fft.fft2(x))), rtol=0, atol=1e-11)
    xr, yr = rng.standard_normal((2, h, w)), rng.standard_normal((2, h, w))
    a, b = O.two_to_spectrum_2d(-1, -1, xr, yr)
    np.testing.assert_allclose(a, np.fft.fft2(xr), rtol=0, atol=1e-11 * h * w)
    np.testing.assert_allclose(b, np.fft.fft2(yr), rtol=0, atol=1e-11 * h * w)
    xs, ys = O.two_from_spectrum_2d(1, 1, a, b)
    np.testing.assert_allclose(xs, h * w * xr, rtol=0, atol=1e-10 * h * w)
    np.testing.assert_allclose(ys, h * w * yr, rtol=0, atol=1e-10 * h * w)


def test_complex_karatsuba_i64_matches_four_products():
    """Convolution.complex karatsuba (Adhoc.hs:114-133) on Int64: the Gauss
    three-product form equals the textbook four products
    (xr*yr - xi*yi, xr*yi + xi*yr) of the schoolbook convolution, in
    wrap-around arithmetic (coefficients up to 2^62), linear and cyclic."""
    def school(a, b):
        return O.schoolbook(a, b)
    for L, m in ((1, 99), (2, 99), (33, 2 ** 15), (64, 2 ** 62), (129, 2 ** 62)):
        r = np.random.default_rng(300 + L)
        xr, xi, yr, yi = (r.integers(-m, m, (3, L)) for _ in range(4))
        re, im = O.complex_karatsuba_i64(xr, xi, yr, yi)
        assert np.array_equal(re, school(xr, yr) - school(xi, yi))
        assert np.array_equal(im, school(xr, yi) + school(xi, yr))
        cre, cim = O.complex_karatsuba_i64(xr, xi, yr, yi, cyclic=True)
        for lin, cyc in ((re, cre), (im, cim)):
            w = lin[:, :L].copy()
            w[:, :L - 1] += lin[:, L:]
            assert np.array_equal(cyc, w)
    # a sign slip in the Gauss recombination fails the comparison
    r = np.random.default_rng(90)
    xr, xi, yr, yi = (r.integers(-99, 99, (1, 8)) for _ in range(4))
    bad_im = O.karatsuba(xr + xi, yr + yi) - O.karatsuba(xr, yr) + O.karatsuba(xi, yi)
    assert not np.array_equal(bad_im, school(xr, yi) + school(xi, yr))
