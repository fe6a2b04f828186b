"""GPU parity of the transform path against the oracle (the C restatement of
the reference's algorithms), through the C ABI.

Tolerances: FP64 relative L2 <= 1e-12 against the FP64 oracle; FP32 relative
L2 <= 1e-6 against the FP64 oracle (the FP32 path's target); Int64 bit-exact.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from accelerate_fourier import _abi, convolution, planned, real  # noqa: E402
from accelerate_fourier.sign import forward, inverse, Sign  # noqa: E402
from accelerate_fourier import synthetic  # noqa: E402

DEV = "cuda"


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def to_dev(x, dtype):
    return torch.from_numpy(np.ascontiguousarray(x)).to(DEV, dtype)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


# lengths covering every device plan: single-pass Stockham (all radices),
# four-step, Rader (single and four-step core) and Bluestein (single, four-step)
LENGTHS_SMALL = list(range(0, 41)) + [48, 60, 64, 77, 97, 100, 121, 127, 128, 169, 255, 256, 257]
LENGTHS_MID = [1000, 1024, 2048, 2310, 4096, 4099, 8191, 8192, 10007, 15015, 16384]
LENGTHS_BIG = [16385 * 0 + 20000, 32768, 65521, 65536, 100003]


def _check_c2c(n, sign, prec, layout, rows=3, seed=7):
    x = synthetic.complex_uniform(seed + n, (rows, n))
    ref = orc_transform(sign, x)
    ct = torch.complex128 if prec == "f64" else torch.complex64
    rt = torch.float64 if prec == "f64" else torch.float32
    t = planned.transform(Sign(sign), n)
    if layout == "planar":
        xd = (to_dev(x.real, rt), to_dev(x.imag, rt))
        y = t(xd)
        got = host(y[0]) + 1j * host(y[1])
    else:
        y = t(to_dev(x, ct))
        got = host(y)
    tol = tol_f64(n) if prec == "f64" else 1e-6
    if n == 0:
        assert got.shape == (rows, 0)
        return
    err = rel_l2(got, ref)
    assert err <= tol, f"n={n} sign={sign} {prec} {layout}: rel-L2 {err:.3e} ({planned.describe(n)})"


def tol_f64(n):
    """The oracle follows the reference's Bluestein chirp, whose phase pi k^2/n
    is not reduced modulo 2n (Private.hs:428-436): its own FP64 error grows
    like eps * n on chirp plans, so the bound widens by that much there."""
    import oracle
    return 1e-12 + (4e-15 * n if oracle.plan_string(n).startswith("chirp") else 0.0)


def orc_transform(sign, x):
    import oracle
    return oracle.transform(sign, x)


@pytest.mark.parametrize("n", LENGTHS_SMALL)
@pytest.mark.parametrize("sign", [-1, 1])
def test_c2c_small_f64(n, sign):
    _check_c2c(n, sign, "f64", "interleaved")


@pytest.mark.parametrize("n", LENGTHS_SMALL[::3])
def test_c2c_small_f32_planar(n):
    _check_c2c(n, -1, "f32", "planar")


@pytest.mark.parametrize("n", LENGTHS_MID)
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_c2c_mid(n, prec):
    _check_c2c(n, -1, prec, "interleaved", rows=2)
    _check_c2c(n, 1, prec, "planar", rows=2)


# products of three or more pairwise coprime radices (2^a, a <= 4, and
# distinct primes up to 13): single passes in prime-factor order; FP64 single
# passes end at 8192
PRIME_FACTOR_LENGTHS = [30, 42, 60, 105, 240, 1155, 3003, 4620, 9240, 15015, 16016]


@pytest.mark.parametrize("n", PRIME_FACTOR_LENGTHS)
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_prime_factor_order(n, prec):
    """Good-Thomas (Planned.hs:584-617) inside the single LDS pass: the
    skew-grid input map and the Chinese-remainder output map applied as LDS
    positions, no twiddles between the coprime factors.  Both signs, both
    layouts, against the oracle (whose planner takes Good-Thomas for the odd
    coprime lengths, Planned.hs:263-270) at the FP64 / FP32 tolerances, and the
    same lengths in Cooley-Tukey order (AF_FLAG_COOLEY_TUKEY) likewise; the two
    orders agree to rounding but not bit for bit, so both paths ran."""
    p = _abi.F64 if prec == "f64" else _abi.F32
    single = n <= (8192 if prec == "f64" else 16384)
    assert planned.describe(n, precision=p).split(":")[1].startswith("primefactor" if single else "fourstep")
    x = synthetic.complex_uniform(300 + n, (3, n))
    rt = torch.float64 if prec == "f64" else torch.float32
    ct = torch.complex128 if prec == "f64" else torch.complex64
    tol = 1e-12 if prec == "f64" else 1e-6
    for sign in (-1, 1):
        ref = orc_transform(sign, x)
        outs = []
        for cooley_tukey in (False, True):
            t = planned.Transform(Sign(sign), n, cooley_tukey=cooley_tukey)
            got_i = host(t(to_dev(x, ct)))
            yp = t((to_dev(x.real, rt), to_dev(x.imag, rt)))
            got_p = host(yp[0]) + 1j * host(yp[1])
            for name, got in (("interleaved", got_i), ("planar", got_p)):
                err = rel_l2(got, ref)
                assert err <= tol, (n, prec, sign, cooley_tukey, name, err)
            outs.append(got_i)
        if single:
            assert not np.array_equal(outs[0], outs[1]), "the two orders gave identical bits"


@pytest.mark.parametrize("n", LENGTHS_BIG)
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_c2c_big(n, prec):
    _check_c2c(n, -1, prec, "interleaved", rows=2)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("layout", ["interleaved", "planar"])
def test_c2c_65536_batch64(prec, layout):
    """64 x 65536 (the batched four-step shape) with every tile resident at
    once, run twice: both runs within tolerance of numpy's FP64 FFT and
    identical to each other."""
    x = synthetic.complex_normal(1, (64, 65536))
    ref = np.fft.fft(x, axis=1)
    ct = torch.complex128 if prec == "f64" else torch.complex64
    rt = torch.float64 if prec == "f64" else torch.float32
    t = planned.transform(forward, 65536)
    outs = []
    for _ in range(2):
        if layout == "planar":
            y = t((to_dev(x.real, rt), to_dev(x.imag, rt)))
            outs.append(host(y[0]) + 1j * host(y[1]))
        else:
            outs.append(host(t(to_dev(x, ct))))
    tol = 1e-13 if prec == "f64" else 1e-6
    for got in outs:
        worst = max(rel_l2(got[r], ref[r]) for r in range(64))
        assert worst <= tol, worst
    assert np.array_equal(outs[0], outs[1])


def test_config1_plain_f64():
    """256 x 1024 Complex Double, U[-1,1) from seed 0: forward spectrum and
    round trip within 1e-12 relative L2 of the FP64 oracle."""
    import oracle
    x = synthetic.complex_uniform(0, (256, 1024))
    xd = to_dev(x, torch.complex128)
    X = planned.transform(forward, 1024)(xd)
    back = planned.transform(inverse, 1024)(X)
    assert rel_l2(host(X), oracle.transform(-1, x)) <= 1e-12
    assert rel_l2(host(back) / 1024, x) <= 1e-12


@pytest.mark.parametrize("n", [2, 4, 6, 8, 10, 16, 30, 64, 100, 1024, 8192, 16384, 32768])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_real_spectrum(n, prec):
    import oracle
    rt = torch.float64 if prec == "f64" else torch.float32
    tol = 1e-12 if prec == "f64" else 1e-6
    x = synthetic.normal(11 + n, 3 * n).reshape(3, n)
    S = real.to_spectrum(planned.transform(forward, n // 2), to_dev(x, rt))
    ref = oracle.to_spectrum(x)
    assert rel_l2(host(S), ref) <= tol
    y = real.from_spectrum(planned.transform(inverse, n // 2), to_dev(ref, S.dtype))
    assert rel_l2(host(y), oracle.from_spectrum(ref)) <= tol


@pytest.mark.parametrize("n", [100, 1024, 2048, 4096, 8192, 16384, 32768])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_real_convolution(n, prec):
    """The fused real convolution and its adjoint against the oracle's
    fromSpectrum(toSpectrum x * toSpectrum k) and the correlations
    fromSpectrum(S_w conj S_k), fromSpectrum(S_w conj S_x): the generic kernel
    (n = 100, and half lengths 8192 and 16384 above the power-of-two one) and
    the power-of-two one at half lengths 512 ... 4096 (the register-twiddle
    form from 2048 on)."""
    import oracle
    rt = torch.float64 if prec == "f64" else torch.float32
    tol = 1e-12 if prec == "f64" else 1e-6
    x = synthetic.normal(3 + n, 4 * n).reshape(4, n)
    k = synthetic.normal(4 + n, 4 * n).reshape(4, n) / np.sqrt(n)
    w = synthetic.normal(5 + n, 4 * n).reshape(4, n)
    y = real.convolve(to_dev(x, rt), to_dev(k, rt))
    Sx, Sk, Sw = oracle.to_spectrum(x), oracle.to_spectrum(k), oracle.to_spectrum(w)
    assert rel_l2(host(y), oracle.from_spectrum(Sx * Sk)) <= tol
    dx, dk = real.convolve_vjp(to_dev(x, rt), to_dev(k, rt), to_dev(w, rt))
    assert rel_l2(host(dx), oracle.from_spectrum(Sw * np.conj(Sk))) <= tol
    assert rel_l2(host(dk), oracle.from_spectrum(Sw * np.conj(Sx))) <= tol


@pytest.mark.parametrize("n", [1, 2, 7, 64, 100, 97, 1024, 20000])
def test_convolve_cyclic(n):
    import oracle
    x = synthetic.complex_uniform(20 + n, (2, n))
    y = synthetic.complex_uniform(30 + n, (2, n))
    z = planned.convolve_cyclic(n)(to_dev(x, torch.complex128), to_dev(y, torch.complex128))
    assert rel_l2(host(z), oracle.convolve_cyclic(x, y)) <= 1e-12


@pytest.mark.parametrize("L", [1, 2, 3, 5, 64, 65, 100, 256, 1000, 4096])
def test_karatsuba_bit_exact(L):
    import oracle
    x = synthetic.integers(40 + L, 3 * L, -2 ** 15, 2 ** 15).reshape(3, L)
    y = synthetic.integers(50 + L, 3 * L, -2 ** 15, 2 ** 15).reshape(3, L)
    z = convolution.karatsuba(to_dev(x, torch.int64), to_dev(y, torch.int64))
    assert np.array_equal(host(z), oracle.schoolbook(x, y))


def test_karatsuba_wraparound():
    import oracle
    L = 777
    x = synthetic.integers(5, L, -2 ** 62, 2 ** 62).reshape(1, L)
    y = synthetic.integers(6, L, -2 ** 62, 2 ** 62).reshape(1, L)
    z = convolution.karatsuba(to_dev(x, torch.int64), to_dev(y, torch.int64))
    assert np.array_equal(host(z), oracle.karatsuba(x, y))


def test_karatsuba_ragged_mixed_magnitudes():
    """Pairs take the exact-FP64 leaf products only when their coefficients are
    small enough (bits(x) + bits(y) + 2d + bits(s) <= 53); ragged pairs of both
    kinds, and ones right at the bound, match the oracle's ring arithmetic bit
    for bit."""
    import oracle
    lens = [4096, 777, 4608, 256, 3000, 2049, 1]
    mags = [2 ** 15, 2 ** 62, 2 ** 14, 2 ** 20, 2 ** 15 + 1, 2 ** 40, 2 ** 63 - 1]
    xs, ys = [], []
    for i, (L, m) in enumerate(zip(lens, mags)):
        xs.append(synthetic.integers(60 + i, L, -m, m))
        ys.append(synthetic.integers(70 + i, L, -m, m))
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    z = convolution.karatsuba_ragged(off, to_dev(np.concatenate(xs), torch.int64),
                                     to_dev(np.concatenate(ys), torch.int64))
    z = host(z)
    for p, L in enumerate(lens):
        ref = oracle.karatsuba(xs[p][None, :], ys[p][None, :])[0]
        got = z[2 * off[p] - p: 2 * off[p] - p + 2 * L - 1]
        assert np.array_equal(got, ref), (p, L, mags[p])


def test_karatsuba_matrix_core_bound():
    """The matrix-core kernel takes a pair when bits(max|x|) + bits(max|y|) +
    2T + bits(Q) <= 51 (its subtree products stay exact in FP64 and convert
    back by the 2^52 + 2^51 bias); the integer kernel takes the others.  Pairs
    of length 1024 (P = 2^4 64: T = 1 top level, Q = 512) with magnitudes of
    20 + 19 bits (51: matrix cores), 21 + 19 bits (52) and 20 + 20 bits (52:
    integer kernel), in one ragged batch with a small pair, match exact Python
    integer products bit for bit."""
    L = 1024
    rng = np.random.default_rng(11)
    cases = [(2 ** 19 + 5, 2 ** 18 + 3), (2 ** 20 + 5, 2 ** 18 + 3), (2 ** 19 + 5, 2 ** 19 + 3), (2 ** 15, 2 ** 15)]
    xs, ys = [], []
    for ax, ay in cases:
        x = rng.integers(-(ax - 1), ax, L)
        y = rng.integers(-(ay - 1), ay, L)
        x[0], y[-1] = ax, -ay  # the largest magnitude of each operand, exactly
        xs.append(x.astype(np.int64))
        ys.append(y.astype(np.int64))
    xs.append(np.array([7, -3], dtype=np.int64))
    ys.append(np.array([-2, 5], dtype=np.int64))
    lens = [len(v) for v in xs]
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    z = host(convolution.karatsuba_ragged(off, to_dev(np.concatenate(xs), torch.int64),
                                          to_dev(np.concatenate(ys), torch.int64)))
    for p, n in enumerate(lens):
        ref = [0] * (2 * n - 1)
        xp, yp = [int(v) for v in xs[p]], [int(v) for v in ys[p]]
        for i in range(n):
            for j in range(n):
                ref[i + j] += xp[i] * yp[j]
        got = z[2 * off[p] - p: 2 * off[p] - p + 2 * n - 1]
        assert [int(v) for v in got] == ref, (p, n)


def test_random_lengths_real_and_karatsuba():
    """Seeded random lengths: toSpectrum / fromSpectrum and the real
    convolution with its adjoint at 24 even lengths up to 60000 in both
    precisions (every kernel path: fused power-of-two, generic, packed beyond
    one pass), and one ragged Karatsuba batch of 48 pairs with lengths up to
    6000 (matrix-core, integer and split-above-the-LDS-limit pairs, both
    coefficient ranges) bit for bit against the oracle.  FP32 results meet the
    oracle's FP64 values at 1e-6.  FP64 results meet numpy's FFT at 1e-12 and
    the oracle at the reference's 1e-10 (Fourier.hs:34-35): for half lengths
    that the reference plans with Bluestein (3412, 3413, ...) its chirp phase
    k^2 / n is not reduced modulo 2n before the sine and cosine, and the oracle,
    which restates it, errs by 5e-13 .. 1e-12 against numpy where this path
    stays near 1e-15."""
    import oracle
    rng = np.random.default_rng(23)
    lens = sorted(set(2 * int(h) for h in rng.integers(1, 30000, 24)))

    def conv_np(a, b, conj_b=False):
        fa, fb = np.fft.rfft(a), np.fft.rfft(b)
        return np.fft.irfft(fa * (np.conj(fb) if conj_b else fb), n=a.shape[-1]) * a.shape[-1]

    for i, n in enumerate(lens):
        x = synthetic.normal(300 + i, 2 * n).reshape(2, n)
        k = synthetic.normal(400 + i, 2 * n).reshape(2, n) / np.sqrt(n)
        w = synthetic.normal(500 + i, 2 * n).reshape(2, n)
        Sx, Sk, Sw = oracle.to_spectrum(x), oracle.to_spectrum(k), oracle.to_spectrum(w)
        orc = [Sx, oracle.from_spectrum(Sx), oracle.from_spectrum(Sx * Sk), oracle.from_spectrum(Sw * np.conj(Sk)),
               oracle.from_spectrum(Sw * np.conj(Sx))]
        npy = [np.fft.fft(x), None, conv_np(x, k), conv_np(w, k, True), conv_np(w, x, True)]
        for rt in (torch.float64, torch.float32):
            S = real.to_spectrum(planned.transform(forward, n // 2), to_dev(x, rt))
            y = real.from_spectrum(planned.transform(inverse, n // 2), to_dev(Sx, S.dtype))
            yc = real.convolve(to_dev(x, rt), to_dev(k, rt))
            dx, dk = real.convolve_vjp(to_dev(x, rt), to_dev(k, rt), to_dev(w, rt))
            for j, got in enumerate(host(t) for t in (S, y, yc, dx, dk)):
                if rt == torch.float32:
                    assert rel_l2(got, orc[j]) <= 1e-6, (n, j, "f32")
                else:
                    assert rel_l2(got, orc[j]) <= 1e-10, (n, j, "f64 vs oracle")
                    ref = npy[j] if npy[j] is not None else np.fft.ifft(Sx, axis=-1).real * n
                    assert rel_l2(got, ref) <= 1e-12, (n, j, "f64 vs numpy")
    klens = [int(v) for v in rng.integers(1, 6001, 48)]
    mags = [2 ** 15 if p % 3 else 2 ** 40 for p in range(len(klens))]
    xs = [synthetic.integers(600 + p, L, -m, m) for p, (L, m) in enumerate(zip(klens, mags))]
    ys = [synthetic.integers(700 + p, L, -m, m) for p, (L, m) in enumerate(zip(klens, mags))]
    off = np.concatenate([[0], np.cumsum(klens)]).astype(np.int64)
    z = host(convolution.karatsuba_ragged(off, to_dev(np.concatenate(xs), torch.int64),
                                          to_dev(np.concatenate(ys), torch.int64)))
    for p, L in enumerate(klens):
        ref = oracle.karatsuba(xs[p][None, :], ys[p][None, :])[0]
        assert np.array_equal(z[2 * off[p] - p: 2 * off[p] - p + 2 * L - 1], ref), (p, L, mags[p])
    print(f"real: {len(lens)} even lengths {lens[0]}..{lens[-1]}; Karatsuba: {len(klens)} pairs, "
          f"{sum(L > 4608 for L in klens)} above the LDS limit")


@pytest.mark.parametrize("L", [4609, 5000, 16384, 100000])
def test_karatsuba_past_lds_limit(L):
    """Convolution.Adhoc.karatsuba has no length limit (Adhoc.hs:33-64): pairs
    whose padded length exceeds one workgroup's LDS split their top levels in
    global memory.  Linear and cyclic, bit for bit against the oracle's
    restatement of the reference's Karatsuba, with coefficients small enough for
    the exact-FP64 leaves and large enough to wrap around."""
    import oracle
    for m in (2 ** 15, 2 ** 62):
        batch = 2 if L < 20000 else 1
        x = synthetic.integers(900 + L + m % 7, batch * L, -m, m).reshape(batch, L)
        y = synthetic.integers(950 + L + m % 7, batch * L, -m, m).reshape(batch, L)
        xd, yd = to_dev(x, torch.int64), to_dev(y, torch.int64)
        ref = oracle.karatsuba(x, y)
        z = host(convolution.karatsuba(xd, yd))
        assert np.array_equal(z, ref), (L, m)
        # Convolution.cyclic: the tail wrapped onto the head (Adhoc.hs:104-111)
        refc = ref[:, :L].copy()
        refc[:, :L - 1] += ref[:, L:]
        zc = host(convolution.cyclic(convolution.karatsuba)(xd, yd))
        assert np.array_equal(zc, refc), (L, m)


def test_karatsuba_ragged_and_complex_past_lds_limit():
    """Ragged pairs on both sides of the LDS limit in one call, and the Gauss
    form of Convolution.complex over pairs past it, against the oracle."""
    import oracle
    lens = [5000, 300, 9000, 4608, 4609, 1, 12345]
    xs = [synthetic.integers(1200 + i, L, -2 ** 20, 2 ** 20) for i, L in enumerate(lens)]
    ys = [synthetic.integers(1300 + i, L, -2 ** 20, 2 ** 20) for i, L in enumerate(lens)]
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    z = host(convolution.karatsuba_ragged(off, to_dev(np.concatenate(xs), torch.int64),
                                          to_dev(np.concatenate(ys), torch.int64)))
    for p, L in enumerate(lens):
        ref = oracle.karatsuba(xs[p][None, :], ys[p][None, :])[0]
        assert np.array_equal(z[2 * off[p] - p: 2 * off[p] - p + 2 * L - 1], ref), (p, L)
    L = 6000
    xr, xi, yr, yi = (synthetic.integers(1400 + i, L, -2 ** 40, 2 ** 40).reshape(1, L) for i in range(4))
    dev = [to_dev(v, torch.int64) for v in (xr, xi, yr, yi)]
    for cyc in (False, True):
        conv = convolution.complex(convolution.cyclic(convolution.karatsuba) if cyc else convolution.karatsuba)
        zr, zi = conv((dev[0], dev[1]), (dev[2], dev[3]))
        rr, ri = oracle.complex_karatsuba_i64(xr, xi, yr, yi, cyclic=cyc)
        assert np.array_equal(host(zr), rr) and np.array_equal(host(zi), ri), cyc


def test_uneven_batch():
    import oracle
    lengths = [64, 1, 0, 60, 97, 257, 1000, 4099, 20000, 65536, 17, 65521, 100003, 5]
    off = np.concatenate([[0], np.cumsum(lengths)])
    x = synthetic.complex_uniform(2, (int(off[-1]),))
    from accelerate_fourier.batch import UnevenBatch
    b = UnevenBatch(lengths, forward, "f64")
    y = host(b(to_dev(x, torch.complex128)))
    for i, n in enumerate(lengths):
        if n == 0:
            continue
        ref = oracle.transform(-1, x[off[i]:off[i + 1]])
        assert rel_l2(y[off[i]:off[i + 1]], ref) <= tol_f64(n), f"row {i} length {n}"


# primes whose Bluestein pad is a power of two from 2^17 to 2^22 (four-step
# cores of power-of-two factors: the extended fast kernels)
BLUESTEIN_P2 = [33013, 70003, 140009, 300007, 600011, 1100009]


@pytest.mark.parametrize("n", BLUESTEIN_P2)
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_bluestein_power_of_two_pad(n, prec):
    """Two rows, both layouts and signs, against numpy's FP64 FFT (which the
    oracle's Bluestein restatement matches on the smaller primes above)."""
    desc = planned.describe(n, -1, _abi.C2C, _abi.F64 if prec == "f64" else _abi.F32)
    m = 1 << (2 * n - 2).bit_length()
    assert f"bluestein{n}(fourstep{m}(" in desc, desc
    x = synthetic.complex_normal(17, (2, n))
    ct = torch.complex128 if prec == "f64" else torch.complex64
    rt = torch.float64 if prec == "f64" else torch.float32
    tol = 1e-11 if prec == "f64" else 1e-6
    for sign, layout in ((-1, "interleaved"), (1, "planar")):
        ref = np.fft.fft(x, axis=1) if sign < 0 else np.fft.ifft(x, axis=1) * n
        t = planned.transform(Sign(sign), n)
        if layout == "planar":
            y = t((to_dev(x.real, rt), to_dev(x.imag, rt)))
            got = host(y[0]) + 1j * host(y[1])
        else:
            got = host(t(to_dev(x, ct)))
        for r in range(2):
            assert rel_l2(got[r], ref[r]) <= tol, (sign, layout, r, rel_l2(got[r], ref[r]))


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_uneven_batch_power_of_two_groups(prec):
    """Rows whose passes run as grouped extended fast launches (Bluestein with
    power-of-two pads, a 65536 four-step) between generic ones, each row
    against numpy's FP64 FFT."""
    lengths = [100003, 600011, 70003, 65536, 0, 300007, 1024, 97, 100003, 1100009]
    off = np.concatenate([[0], np.cumsum(lengths)])
    x = synthetic.complex_normal(19, (int(off[-1]),))
    from accelerate_fourier.batch import UnevenBatch
    ct = torch.complex128 if prec == "f64" else torch.complex64
    b = UnevenBatch(lengths, forward, prec)
    xd = to_dev(x, ct)
    y1 = host(b(xd))
    y2 = host(b(xd))
    assert np.array_equal(y1, y2)
    tol = 1e-11 if prec == "f64" else 1e-6
    for i, n in enumerate(lengths):
        if n == 0:
            continue
        ref = np.fft.fft(x[off[i]:off[i + 1]])
        assert rel_l2(y1[off[i]:off[i + 1]], ref) <= tol, f"row {i} length {n}"


@pytest.mark.parametrize("n", [86, 58, 40000, 2 * 65521])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_real_spectrum_packed(n, prec):
    """Half lengths that are prime or above one LDS pass: half-length C2C plan
    plus the packing kernels."""
    import oracle
    rt = torch.float64 if prec == "f64" else torch.float32
    x = synthetic.normal(13 + n, 2 * n).reshape(2, n)
    S = real.to_spectrum(planned.transform(forward, n // 2), to_dev(x, rt))
    ref = oracle.to_spectrum(x)
    tol = tol_f64(n // 2) if prec == "f64" else 1e-6
    assert rel_l2(host(S), ref) <= tol
    y = real.from_spectrum(planned.transform(inverse, n // 2), to_dev(ref, S.dtype))
    assert rel_l2(host(y), oracle.from_spectrum(ref)) <= tol
    from accelerate_fourier import autograd as AG  # noqa: F401
    g = to_dev(synthetic.complex_uniform(5, (2, n)), torch.complex128 if prec == "f64" else torch.complex64)
    dx = host(real.vjp_to_spectrum(g))
    gg = host(g).astype(np.complex128)
    ref_dx = np.real(np.fft.ifft(gg) * n)  # Re(F^H g)
    assert rel_l2(dx, ref_dx) <= (1e-11 if prec == "f64" else 1e-6)


@pytest.mark.parametrize("n", [86, 40000])
def test_real_convolution_packed(n):
    import oracle
    x = synthetic.normal(3, 2 * n).reshape(2, n)
    k = synthetic.normal(4, 2 * n).reshape(2, n) / np.sqrt(n)
    w = synthetic.normal(5, 2 * n).reshape(2, n)
    xd, kd, wd = to_dev(x, torch.float64), to_dev(k, torch.float64), to_dev(w, torch.float64)
    y = real.convolve(xd, kd)
    Sx, Sk, Sw = oracle.to_spectrum(x), oracle.to_spectrum(k), oracle.to_spectrum(w)
    assert rel_l2(host(y), oracle.from_spectrum(Sx * Sk)) <= tol_f64(n // 2)
    dx, dk = real.convolve_vjp(xd, kd, wd)
    assert rel_l2(host(dx), oracle.from_spectrum(Sw * np.conj(Sk))) <= tol_f64(n // 2)
    assert rel_l2(host(dk), oracle.from_spectrum(Sw * np.conj(Sx))) <= tol_f64(n // 2)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("layout", ["planar", "interleaved"])
@pytest.mark.parametrize("sign", [-1, 1])
def test_three_pass_2p24(prec, layout, sign):
    """n = 2^24 runs as three passes of 256-point transforms (the last one per
    transform of the batch); two rows against numpy's FP64 FFT, which the
    oracle's restatement matches on the smaller lengths above."""
    n, rows = 1 << 24, 2
    assert planned.describe(n, sign, _abi.C2C, _abi.F64 if prec == "f64" else _abi.F32).split(":")[1].startswith(
        "threepass16777216")
    x = synthetic.complex_normal(11, (rows, n))
    ref = np.fft.fft(x, axis=1) if sign < 0 else np.fft.ifft(x, axis=1) * n
    ct = torch.complex128 if prec == "f64" else torch.complex64
    rt = torch.float64 if prec == "f64" else torch.float32
    t = planned.transform(Sign(sign), n)
    if layout == "planar":
        y = t((to_dev(x.real, rt), to_dev(x.imag, rt)))
        got = host(y[0]) + 1j * host(y[1])
    else:
        got = host(t(to_dev(x, ct)))
    tol = 1e-13 if prec == "f64" else 1e-6
    for r in range(rows):
        assert rel_l2(got[r], ref[r]) <= tol, (r, rel_l2(got[r], ref[r]))


@pytest.mark.parametrize("n,batch,layout,prec", [(65536, 8, "planar", "f32"), (20000, 3, "interleaved", "f32"),
                                                (100003, 2, "planar", "f32"), (1000, 16, "interleaved", "f64"),
                                                # prime-factor order: the tuner's other radix orders run the
                                                # twiddled order instead
                                                (15015, 4, "interleaved", "f32"), (1155, 8, "planar", "f64")])
def test_autotuned_transform(n, batch, layout, prec):
    """After af_plan_tune the transform (and its adjoint) with the tuned shapes
    still matches numpy's FP64 FFT; the wisdom text round-trips."""
    t = planned.transform(forward, n)
    report = t.tune(batch, prec, layout)
    assert report.count("pass ") >= 1, report
    x = synthetic.complex_normal(23, (batch, n))
    ct = torch.complex128 if prec == "f64" else torch.complex64
    rt = torch.float64 if prec == "f64" else torch.float32
    tol = 1e-11 if prec == "f64" else 1e-6
    if layout == "planar":
        y = t((to_dev(x.real, rt), to_dev(x.imag, rt)))
        got = host(y[0]) + 1j * host(y[1])
        g = t.vjp((to_dev(x.real, rt), to_dev(x.imag, rt)))
        gv = host(g[0]) + 1j * host(g[1])
    else:
        got = host(t(to_dev(x, ct)))
        gv = host(t.vjp(to_dev(x, ct)))
    assert rel_l2(got, np.fft.fft(x, axis=1)) <= tol, report
    assert rel_l2(gv, np.fft.ifft(x, axis=1) * n) <= tol, report
    text = planned.wisdom_export()
    key = f"{prec} 0 {n} {batch} {1 if layout == 'planar' else 0} :"
    assert key in text
    planned.wisdom_import(text)
    assert planned.wisdom_export() == text


@pytest.mark.parametrize("n", [1, 2, 7, 64, 97, 1000, 4099, 65536, 100003])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_two_for_one_real_packing(n, prec):
    """twoToSpectrum / twoFromSpectrum (Real.hs:129-196) against the oracle,
    planar and interleaved spectra, and the reference's properties
    (Fourier.hs:503-536): the two spectra are those of x and y, and the inverse
    transform of them gives back n x and n y."""
    import oracle
    rt = torch.float64 if prec == "f64" else torch.float32
    ct = torch.complex128 if prec == "f64" else torch.complex64
    tol = tol_f64(n) if prec == "f64" else 1e-6
    x = synthetic.normal(31 + n, 3 * n).reshape(3, n)
    y = synthetic.normal(41 + n, 3 * n).reshape(3, n)
    ra, rb = oracle.two_to_spectrum(-1, x, y)
    a, b = real.two_to_spectrum(planned.transform(forward, n), to_dev(x, rt), to_dev(y, rt))
    assert rel_l2(host(a), ra) <= tol and rel_l2(host(b), rb) <= tol
    ap = (torch.empty(3, n, dtype=rt, device=DEV), torch.empty(3, n, dtype=rt, device=DEV))
    bp = (torch.empty(3, n, dtype=rt, device=DEV), torch.empty(3, n, dtype=rt, device=DEV))
    real.two_to_spectrum(planned.transform(forward, n), to_dev(x, rt), to_dev(y, rt), out=(ap, bp))
    assert rel_l2(host(ap[0]) + 1j * host(ap[1]), ra) <= tol
    assert rel_l2(host(bp[0]) + 1j * host(bp[1]), rb) <= tol
    xs, ys = real.two_from_spectrum(planned.transform(inverse, n), to_dev(ra, ct), to_dev(rb, ct))
    ox, oy = oracle.two_from_spectrum(1, ra, rb)
    assert rel_l2(host(xs), ox) <= tol and rel_l2(host(ys), oy) <= tol
    assert rel_l2(host(xs), n * x) <= 10 * tol and rel_l2(host(ys), n * y) <= 10 * tol


def test_untangle_entangle_elementwise():
    """untangleSpectra / entangleSpectra alone: the coefficient formulas of
    Real.hs:182-187 and 231-236 on arbitrary complex data (FP64 exact up to one
    rounding), and untangle . entangle = id on spectra of real signals."""
    n = 1001
    rng = np.random.default_rng(5)
    z = rng.standard_normal((4, n)) + 1j * rng.standard_normal((4, n))
    a, b = real.untangle_spectra(to_dev(z, torch.complex128))
    zr = np.conj(z[:, (-np.arange(n)) % n])
    assert rel_l2(host(a), (z + zr) / 2) <= 1e-15 and rel_l2(host(b), (z - zr) * (-0.5j)) <= 1e-15
    sa, sb = np.fft.fft(rng.standard_normal((4, n))), np.fft.fft(rng.standard_normal((4, n)))
    e = real.entangle_spectra(to_dev(sa, torch.complex128), to_dev(sb, torch.complex128))
    assert rel_l2(host(e), sa + 1j * sb) <= 1e-15
    a2, b2 = real.untangle_spectra(e)
    assert rel_l2(host(a2), sa) <= 1e-14 and rel_l2(host(b2), sb) <= 1e-14


@pytest.mark.parametrize("h,w", [(1, 1), (7, 5), (64, 48), (256, 1024), (1000, 3), (4099, 16), (16, 4099),
                                 (65536, 2), (3, 65536)])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_transform_2d(h, w, prec):
    """Preprocessed.transform2d (Preprocessed.hs:141-148) against the oracle's
    composition of 1-D transforms, both layouts, and the reference's property
    transpose (transform2d (tw, th) x) = transform2d (th, tw) (transpose x)
    (Fourier.hs:537-551) on the device."""
    import oracle
    ct = torch.complex128 if prec == "f64" else torch.complex64
    rt = torch.float64 if prec == "f64" else torch.float32
    tol = 1e-12 if prec == "f64" else 2e-6
    x = synthetic.complex_uniform(h * 7 + w, (2, h, w))
    ref = oracle.transform2d(-1, 1, x)
    t = planned.transform2d(planned.transform(forward, w), planned.transform(inverse, h))
    got = host(t(to_dev(x, ct)))
    assert rel_l2(got, ref) <= tol
    yp = t((to_dev(x.real, rt), to_dev(x.imag, rt)))
    assert rel_l2(host(yp[0]) + 1j * host(yp[1]), ref) <= tol
    tv = planned.transform2d(planned.transform(inverse, h), planned.transform(forward, w))
    lhs = host(planned.transpose(t(to_dev(x, ct))))
    rhs = host(tv(planned.transpose(to_dev(x, ct))))
    assert rel_l2(lhs, rhs) <= tol


@pytest.mark.parametrize("k,m,n", [(1, 1, 1), (2, 3, 5), (5, 12, 16), (3, 4099, 2), (64, 32, 48), (7, 2, 1000)])
def test_transform_3d(k, m, n):
    """Preprocessed.transform3d (Preprocessed.hs:150-161) against the oracle and
    the property cycleDim3 (transform3d (tw, th, td) x) =
    transform3d (th, td, tw) (cycleDim3 x) (Fourier.hs:560-575)."""
    import oracle
    x = synthetic.complex_uniform(k + m + n, (2, k, m, n))
    for prec, ct, tol in (("f64", torch.complex128, 1e-12), ("f32", torch.complex64, 2e-6)):
        t = planned.transform3d(planned.transform(forward, n), planned.transform(inverse, m),
                                planned.transform(forward, k))
        got = host(t(to_dev(x, ct)))
        assert rel_l2(got, oracle.transform3d(-1, 1, -1, x)) <= tol, prec
        tv = planned.transform3d(planned.transform(inverse, m), planned.transform(forward, k),
                                 planned.transform(forward, n))
        lhs = host(planned.cycle_dim3(t(to_dev(x, ct))))
        rhs = host(tv(planned.cycle_dim3(to_dev(x, ct))))
        assert rel_l2(lhs, rhs) <= tol, prec


def test_random_shapes_2d_3d():
    """Seeded random 2-D and 3-D shapes (extents 1 .. 3000, mixed signs per
    axis, both precisions) against the oracle's composition of 1-D transforms:
    every axis plan kind (power of two, smooth, prime) as a strided pass or
    through the transpose.  The FP64 bound widens by 4e-15 per point of the
    longer axis: the oracle restates the reference's Bluestein, whose chirp
    phase is not reduced, so its own error grows with the length."""
    import oracle
    rng = np.random.default_rng(31)
    for i in range(12):
        h, w = (int(v) for v in rng.integers(1, 3000, 2))
        if h * w > 1 << 21:
            w = max(1, (1 << 21) // h)
        x = synthetic.complex_uniform(800 + i, (1, h, w))
        sh, sw = (int(v) for v in rng.choice([-1, 1], 2))
        ref = oracle.transform2d(sw, sh, x)
        for ct, tol in ((torch.complex128, max(1e-12, 4e-15 * max(h, w))), (torch.complex64, 2e-6)):
            t = planned.transform2d(planned.transform(Sign(sw), w), planned.transform(Sign(sh), h))
            assert rel_l2(host(t(to_dev(x, ct))), ref) <= tol, (h, w, sh, sw, str(ct))
    for i in range(8):
        k, m, n = (int(v) for v in rng.integers(1, 160, 3))
        x = synthetic.complex_uniform(900 + i, (1, k, m, n))
        sk, sm, sn = (int(v) for v in rng.choice([-1, 1], 3))
        ref = oracle.transform3d(sn, sm, sk, x)
        for ct, tol in ((torch.complex128, 1e-12), (torch.complex64, 2e-6)):
            t = planned.transform3d(planned.transform(Sign(sn), n), planned.transform(Sign(sm), m),
                                    planned.transform(Sign(sk), k))
            assert rel_l2(host(t(to_dev(x, ct))), ref) <= tol, (k, m, n, sk, sm, sn, str(ct))


def test_transpose_and_cycle_dim3_exact():
    """The data movements themselves, bit for bit."""
    x = synthetic.complex_uniform(3, (3, 37, 70))
    for ct in (torch.complex128, torch.complex64):
        xd = to_dev(x, ct)
        assert np.array_equal(host(planned.transpose(xd)), np.swapaxes(host(xd), -1, -2))
        x3 = xd.reshape(3, 37, 7, 10)
        assert np.array_equal(host(planned.cycle_dim3(x3)), np.moveaxis(host(x3), -1, -3))


@pytest.mark.parametrize("h,w", [(1, 1), (6, 10), (9, 7), (64, 1000)])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_two_for_one_2d(h, w, prec):
    """twoToSpectrum2d / twoFromSpectrum2d / untangleSpectra2d (Real.hs:138-206)
    against the oracle."""
    import oracle
    rt = torch.float64 if prec == "f64" else torch.float32
    ct = torch.complex128 if prec == "f64" else torch.complex64
    tol = 1e-12 if prec == "f64" else 2e-6
    x = synthetic.normal(h + w, 2 * h * w).reshape(2, h, w)
    y = synthetic.normal(h * w + 1, 2 * h * w).reshape(2, h, w)
    ra, rb = oracle.two_to_spectrum_2d(-1, -1, x, y)
    tw, th = planned.transform(forward, w), planned.transform(forward, h)
    a, b = real.two_to_spectrum_2d(tw, th, to_dev(x, rt), to_dev(y, rt))
    assert rel_l2(host(a), ra) <= tol and rel_l2(host(b), rb) <= tol
    iw, ih = planned.transform(inverse, w), planned.transform(inverse, h)
    xs, ys = real.two_from_spectrum_2d(iw, ih, to_dev(ra, ct), to_dev(rb, ct))
    ox, oy = oracle.two_from_spectrum_2d(1, 1, ra, rb)
    assert rel_l2(host(xs), ox) <= tol and rel_l2(host(ys), oy) <= tol
    z = synthetic.complex_uniform(h + 2 * w, (2, h, w))
    ua, ub = real.untangle_spectra_2d(to_dev(z, ct))
    oa, ob = oracle.untangle_spectra_2d(z)
    assert rel_l2(host(ua), oa) <= tol and rel_l2(host(ub), ob) <= tol


@pytest.mark.parametrize("n", [12, 97, 1000, 4099, 65521])
def test_cache_duplex_and_planners(n):
    """cacheDuplex (Planned.hs:331-341): the inverse of the pair runs as the
    adjoint of the forward plan and equals the inverse transform; every planner
    of transformWithPlanner (Planned.hs:121-130) gives the DFT."""
    x = synthetic.complex_uniform(n + 3, (3, n))
    for sign in (-1, 1):
        ref = orc_transform(sign, x)
        for prec, ct in (("f64", torch.complex128), ("f32", torch.complex64)):
            tol = tol_f64(n) if prec == "f64" else 1e-6
            ci, cf = planned.cache_duplex(n)
            t = planned.transform_with_cache(cf if sign == -1 else ci)
            assert rel_l2(host(t(to_dev(x, ct))), ref) <= tol
            for pl in (planned.plan_with_map_update, planned.plan_decompose_with_map_update,
                       planned.plan_chirp_with_map_update(planned.ceiling_power_of_two),
                       planned.plan_chirp_with_map_update(planned.ceiling_5_smooth)):
                tp = planned.transform_with_planner(pl, Sign(sign), n)
                got = host(tp(to_dev(x, ct)))
                assert rel_l2(got, ref) <= (tol + 4e-15 * n if prec == "f64" else 1e-6), pl


@pytest.mark.parametrize("n", [34, 47, 94, 289, 1009, 4418, 10007, 100003])
def test_transform_decompose_nested(n):
    """Planned.transformDecompose (Planned.hs:114-119) never pads: lengths whose
    prime factors no single device plan serves run as host-composed
    decompositions -- Rader whose p - 1 convolution is itself decomposed
    (47 = Rader(46 = GT(2, Rader(23)))), Good-Thomas over the reference's
    skew-grid maps (34, 94, 4418), Cooley-Tukey for prime powers (289 = 17^2).
    Against the oracle's own decomposition plan, both signs, both precisions,
    both layouts."""
    import oracle
    assert "bluestein" not in planned.describe(n, -1, _abi.C2C, _abi.F64, _abi.ALG_DECOMPOSE)
    x = synthetic.complex_uniform(300 + n, (3, n))
    for sign in (-1, 1):
        ref = oracle.transform(sign, x, algorithm=oracle.DECOMPOSE)
        t = planned.transform_decompose(Sign(sign), n)
        got = host(t(to_dev(x, torch.complex128)))
        assert rel_l2(got, ref) <= 1e-12, (n, sign, "f64")
        y = t((to_dev(x.real, torch.float32), to_dev(x.imag, torch.float32)))
        got32 = host(y[0]) + 1j * host(y[1])
        assert rel_l2(got32, ref) <= 1e-6, (n, sign, "f32")
    # the adjoint runs the plan of the opposite sign (its own decomposition)
    g = synthetic.complex_uniform(400 + n, (2, n))
    vj = host(planned.transform_decompose(forward, n).vjp(to_dev(g, torch.complex128)))
    assert rel_l2(vj, oracle.transform(1, g, algorithm=oracle.DECOMPOSE)) <= 1e-12


@pytest.mark.parametrize("L,m", [(1, 2 ** 15), (2, 2 ** 15), (65, 2 ** 15), (777, 2 ** 62), (4096, 2 ** 15)])
def test_karatsuba_complex_bit_exact(L, m):
    """Convolution.complex karatsuba and Convolution.complex (cyclic karatsuba)
    (Adhoc.hs:104-133) on Int64 (re, im) pairs, bit for bit against the oracle."""
    import oracle
    xr, xi, yr, yi = (synthetic.integers(500 + 4 * L + i, 2 * L, -m, m).reshape(2, L) for i in range(4))
    dev = [to_dev(v, torch.int64) for v in (xr, xi, yr, yi)]
    for cyc in (False, True):
        conv = convolution.complex(convolution.cyclic(convolution.karatsuba) if cyc else convolution.karatsuba)
        zr, zi = conv((dev[0], dev[1]), (dev[2], dev[3]))
        rr, ri = oracle.complex_karatsuba_i64(xr, xi, yr, yi, cyclic=cyc)
        assert np.array_equal(host(zr), rr) and np.array_equal(host(zi), ri), cyc


def test_karatsuba_pairs_by_fp32_fft_convolution():
    """The Karatsuba pairs cast to FP32 and convolved by the zero-padded real
    FFT convolution (real.convolve at M = the next power of two >= 2L-1, inputs
    scaled by 1/M, exact): for small coefficients rounding recovers the exact
    Int64 product; for the 2^15 coefficients of the benchmark the max-abs error
    stays within eps_32 * log2(M) * |x|_2 |y|_2."""
    import oracle
    for m, lens in ((8, [256, 300, 1000, 1024]), (2 ** 15, [256, 2049, 4096])):
        for i, L in enumerate(lens):
            x = synthetic.integers(700 + i, L, -m, m)
            y = synthetic.integers(800 + i, L, -m, m)
            M = 1 << int(np.ceil(np.log2(2 * L - 1)))
            xp = np.zeros((1, M), np.float32)
            yp = np.zeros((1, M), np.float32)
            xp[0, :L] = x.astype(np.float32) / M
            yp[0, :L] = y.astype(np.float32)
            got = host(real.convolve(to_dev(xp, torch.float32), to_dev(yp, torch.float32)))[0, :2 * L - 1]
            exact = oracle.schoolbook(x[None, :], y[None, :])[0]
            err = np.max(np.abs(got.astype(np.float64) - exact))
            bound = 2.0 ** -23 * np.log2(M) * np.linalg.norm(x) * np.linalg.norm(y)
            assert err <= bound, (m, L, err, bound)
            if m == 8:
                assert np.array_equal(np.rint(got).astype(np.int64), exact), (L, err)


def test_karatsuba_full_config_bit_exact():
    """The full Karatsuba workload: 1024 ragged pairs, lengths U[256, 4096]
    (seed 4), coefficients U[-2^15, 2^15) (seeds 40, 41), bit for bit against
    the oracle's ragged Karatsuba."""
    import oracle
    lens = synthetic.integers(4, 1024, 256, 4097)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    x = synthetic.integers(40, int(off[-1]), -2 ** 15, 2 ** 15)
    y = synthetic.integers(41, int(off[-1]), -2 ** 15, 2 ** 15)
    z = host(convolution.karatsuba_ragged(off, to_dev(x, torch.int64), to_dev(y, torch.int64)))
    assert np.array_equal(z, oracle.karatsuba_ragged(off, x, y))


def _lsr(z, k):
    """Logical right shift of int64 words (torch shifts arithmetically)."""
    return (z >> k) & ((1 << (64 - k)) - 1)


def _i64(c):
    return c - (1 << 64) if c >= 1 << 63 else c


def dev_uniform(seed, count, start=0, chunk=1 << 26):
    """synthetic.uniform(seed, count, -1, 1, start) computed on the GPU in FP64:
    the same counter-based splitmix64 stream in int64 words (products wrap)."""
    out = torch.empty(count, dtype=torch.float64, device=DEV)
    g, c1, c2 = _i64(0x9E3779B97F4A7C15), _i64(0xBF58476D1CE4E5B9), _i64(0x94D049BB133111EB)
    for lo in range(0, count, chunk):
        hi = min(count, lo + chunk)
        z = torch.arange(start + lo + 1, start + hi + 1, dtype=torch.int64, device=DEV) * g + seed
        z = (z ^ _lsr(z, 30)) * c1
        z = (z ^ _lsr(z, 27)) * c2
        z = z ^ _lsr(z, 31)
        out[lo:hi] = _lsr(z, 11).to(torch.float64) * (2.0 / 9007199254740992.0) - 1.0
    return out


def _uneven_inputs():
    """The 4096-row batch's lengths and its input, U[-1, 1) from splitmix64
    seed 2 (real parts, then imaginary parts), checked against the host stream."""
    lens, cls = synthetic.uneven_lengths(2)
    tot = int(lens.sum())
    probe = dev_uniform(2, 4096, start=tot - 1000).cpu().numpy()
    assert np.array_equal(probe, synthetic.uniform(2, 4096, start=tot - 1000))
    x64 = torch.complex(dev_uniform(2, tot), dev_uniform(2, tot, start=tot))
    return lens, cls, x64


def _launch_keys(desc):
    """(key text, first row) of every grouped launch in a batch description
    ("launch k chunk c phase p <kernel key> rows=r first=f"); the key text is
    the phase and the kernel key."""
    keys = []
    for line in desc.split("\n")[1:]:
        head, first = line.rsplit(" first=", 1)
        keys.append((head.split(" ", 4)[4].rsplit(" rows=", 1)[0], int(first)))
    return keys


def test_uneven_batch_full_config():
    """The full uneven batch (4096 rows, four length classes, splitmix64 seed 2),
    FP32, in both layouts.  FP32 rows against numpy's FP64 FFT of the same
    (FP32-rounded) inputs -- an implementation independent of this library --
    at 1e-6 relative L2: 32 rows per class spread over its lengths, plus the
    first row of every grouped launch of the plan (each extended-kernel key
    (L, B, conv, sign) and each generic thread class of every phase).  Every
    row of the FP32 batch within 1e-6 of the FP64 batch, and two rows per class
    of the FP64 batch against the oracle."""
    import oracle
    from accelerate_fourier.batch import UnevenBatch
    lens, cls, x64 = _uneven_inputs()
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    b32 = UnevenBatch(lens, forward, "f32")
    keys = _launch_keys(b32.describe())
    assert len(keys) >= 10
    rows = {r for _, r in keys}
    for c in range(len(synthetic.UNEVEN_CLASSES)):
        members = np.nonzero(cls == c)[0]
        members = members[np.argsort(lens[members], kind="stable")]
        rows.update(int(members[i]) for i in np.linspace(0, len(members) - 1, 32).round().astype(int))
    x32 = x64.to(torch.complex64)
    y_int = b32(x32)
    y_pla = b32((x32.real.contiguous(), x32.imag.contiguous()))
    torch.cuda.synchronize()
    worst = {}
    for r in sorted(rows):
        lo, hi = int(off[r]), int(off[r + 1])
        xr = host(x32[lo:hi]).astype(np.complex128)
        ref = np.fft.fft(xr)
        for name, got in (("interleaved", host(y_int[lo:hi])),
                          ("planar", host(y_pla[0][lo:hi]) + 1j * host(y_pla[1][lo:hi]))):
            e = rel_l2(got, ref)
            assert e <= 1e-6, (name, r, int(lens[r]), e)
            c = int(cls[r])
            worst[c] = max(worst.get(c, 0.0), e)
    print(f"uneven batch: {len(rows)} rows vs numpy FP64 in both layouts, launch keys covered: "
          + "; ".join(sorted({k for k, _ in keys})) + "; worst rel-L2 per class " + str(worst))
    del y_pla
    y64 = UnevenBatch(lens, forward, "f64")(x64)
    sizes = [int(n) for n in lens]
    err = y_int.to(torch.complex128) - y64
    d = torch.stack([torch.linalg.vector_norm(r) for r in torch.split(err, sizes)])
    nrm = torch.stack([torch.linalg.vector_norm(r) for r in torch.split(y64, sizes)])
    rel = (d / nrm).cpu().numpy()
    del err
    assert rel.max() <= 1e-6, (int(rel.argmax()), float(rel.max()))
    for c in range(len(synthetic.UNEVEN_CLASSES)):
        members = np.nonzero(cls == c)[0]
        for r in (members[np.argmin(lens[members])], members[np.argmax(lens[members])]):
            n = int(lens[r])
            ref = oracle.transform(-1, host(x64[off[r]:off[r + 1]]))
            got = host(y64[off[r]:off[r + 1]])
            assert rel_l2(got, ref) <= tol_f64(n), (c, int(r), n)
    del x64, y64, y_int, x32


def test_uneven_batch_chunks_and_tuner():
    """The 4096-row FP32 batch with its rows that need intermediates cut into
    chunks of 32 MiB and 160 MiB (every chunk's passes on one shared scratch
    region, rows reordered by shape), and the rows without intermediates on a
    second branch of the graph (160 MiB and unchunked): the same results as the
    default graph
    on the seed-2 inputs, up to the rounding of kernels that a chunk's grouping
    may select differently (1e-6 relative per row, and bit-identical rows for
    the power-of-two groups, whose kernel key does not depend on the grouping);
    then af_batch_tune's report names every candidate and its choice, and the
    tuned graph gives the unchunked results again."""
    from accelerate_fourier.batch import UnevenBatch
    lens, cls, x64 = _uneven_inputs()
    x32 = x64.to(torch.complex64)
    del x64
    sizes = [int(n) for n in lens]
    b = UnevenBatch(lens, forward, "f32")
    ref = b(x32)
    torch.cuda.synchronize()
    for cb, lanes in ((32 << 20, 1), (160 << 20, 2), (0, 3), (0, 2)):
        b.set_schedule(cb, lanes)
        head = b.describe().split("\n")[0]
        assert f"chunk_bytes={cb}" in head and f"lanes={lanes}" in head, head
        nchunks = int(head.split(" chunks=")[1].split()[0])
        assert nchunks > 4 or cb == 0, head
        got = b(x32)
        torch.cuda.synchronize()
        d = torch.stack([torch.linalg.vector_norm(r) for r in torch.split(got - ref, sizes)])
        nrm = torch.stack([torch.linalg.vector_norm(r) for r in torch.split(ref, sizes)])
        rel = (d / nrm).cpu().numpy()
        assert rel.max() <= 1e-6, (cb, float(rel.max()), int(rel.argmax()))
        print(f"chunk_bytes={cb} lanes={lanes}: {nchunks} chunks, max rel-L2 vs unchunked {rel.max():.2e}, "
              f"identical rows {int((rel == 0).sum())} of {len(rel)}")
        del got
    rep = b.tune()
    assert rep.count("chunk_bytes=") == 9 and "chosen chunk_bytes=" in rep, rep
    assert " order=" in rep, rep
    print(rep)
    got = b(x32)
    torch.cuda.synchronize()
    d = torch.stack([torch.linalg.vector_norm(r) for r in torch.split(got - ref, sizes)])
    nrm = torch.stack([torch.linalg.vector_norm(r) for r in torch.split(ref, sizes)])
    assert float((d / nrm).max()) <= 1e-6
    with pytest.raises(ValueError):
        b.set_schedule(-1)
    with pytest.raises(ValueError):
        b.set_schedule(0, 4)
    with pytest.raises(ValueError):
        b.set_schedule(32 << 20, 3)


def test_uneven_batch_two_streams_and_threads():
    """One batch called from two streams back to back (same buffers, then new
    output buffers) and from two host threads: every call's result equals the
    single-stream result (calls on one batch are ordered by the batch)."""
    import threading
    from accelerate_fourier.batch import UnevenBatch
    lengths = [100003, 65536, 4099, 20000, 97, 600011]
    x = to_dev(synthetic.complex_uniform(31, (sum(lengths),)), torch.complex64)
    b = UnevenBatch(lengths, forward, "f32")
    ref = host(b(x))
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    outs = [torch.empty_like(x) for _ in range(4)]
    with torch.cuda.stream(s1):
        b(x, out=outs[0])
    with torch.cuda.stream(s2):
        b(x, out=outs[1])
        b(x, out=outs[0])
    with torch.cuda.stream(s1):
        b(x, out=outs[2])
    torch.cuda.synchronize()
    for o in outs[:3]:
        assert np.array_equal(host(o), ref)

    def work(o):
        st = torch.cuda.Stream()
        with torch.cuda.stream(st):
            for _ in range(3):
                b(x, out=o)
        st.synchronize()
    th = [threading.Thread(target=work, args=(outs[i],)) for i in (1, 3)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    torch.cuda.synchronize()
    assert np.array_equal(host(outs[1]), ref) and np.array_equal(host(outs[3]), ref)


def test_concurrent_plans_and_executes():
    """Plans are immutable after creation and creation is thread-safe: plans
    made concurrently from host threads (the C ABI releases the interpreter
    during the call), and one plan executed concurrently from four threads,
    each on its own stream, give every caller its own correct result (plans
    with workspace: four-step 65536, Bluestein 100003, Rader 4099)."""
    import ctypes
    import threading
    import oracle

    handles, errs = [None] * 6, []
    lens = [65536, 100003, 4099, 1000, 65521, 77]

    def create(i):
        try:
            d = _abi.PlanDesc(lens[i], -1, _abi.C2C, _abi.F32, _abi.ALG_AUTO, 0)
            h = ctypes.c_void_p()
            _abi.check(_abi.lib().af_plan_create(ctypes.byref(d), ctypes.byref(h)), "af_plan_create")
            handles[i] = h
        except Exception as e:  # reported below
            errs.append(e)
    th = [threading.Thread(target=create, args=(i,)) for i in range(len(lens))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs, errs
    for i, n in enumerate(lens):
        x = synthetic.complex_uniform(950 + i, (2, n))
        xd = to_dev(x, torch.complex64)
        y = torch.empty_like(xd)
        re, im, _ = _arrays_view(xd)
        ore, oim, _ = _arrays_view(y)
        _abi.check(_abi.lib().af_execute(handles[i], 2, re, im, ore, oim, None, None), "af_execute")
        assert rel_l2(host(y), orc_transform(-1, x)) <= 1e-6, n
        _abi.lib().af_plan_destroy(handles[i])

    for n in (65536, 100003, 4099):
        t = planned.transform(forward, n)
        xs = [synthetic.complex_uniform(900 + i, (3, n)) for i in range(4)]
        t(to_dev(xs[0], torch.complex64))
        outs, errs = [None] * 4, []

        def work(i):
            try:
                s = torch.cuda.Stream()
                with torch.cuda.stream(s):
                    xd = torch.from_numpy(xs[i]).to(DEV, torch.complex64)
                    for _ in range(3):
                        y = t(xd)
                    s.synchronize()
                    outs[i] = y.cpu().numpy()
            except Exception as e:  # reported below
                errs.append(e)
        th = [threading.Thread(target=work, args=(i,)) for i in range(4)]
        for w in th:
            w.start()
        for w in th:
            w.join()
        assert not errs, errs
        for i in range(4):
            assert rel_l2(outs[i], orc_transform(-1, xs[i])) <= 1e-6, (n, i)


def _arrays_view(x):
    from accelerate_fourier import _arrays
    return _arrays.complex_view(x)
