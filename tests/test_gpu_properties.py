"""The reference's QuickCheck properties (test/Test/Data/Array/Accelerate/
Fourier.hs:298-587) ported to the GPU path: Complex Double inputs uniform in
[-1,1]^2, lengths 0..99 drawn as QuickCheck's listOf does, absolute tolerance
1e-10 (Fourier.hs:34-35).  Cross-implementation properties compare the plans
the reference compares (generic vs chirp2 / chirp235 / decompose)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from accelerate_fourier import convolution, planned, real, synthetic  # noqa: E402
from accelerate_fourier.sign import Sign, forward, inverse  # noqa: E402
from accelerate_fourier.utility import scale_down  # noqa: E402

TOL = 1e-10
RNG = np.random.default_rng(2024)
LENS = sorted(set([0, 1, 2, 3, 5, 7, 13, 17, 31, 47, 64, 97] + list(RNG.integers(0, 100, 24))))


def c(n, seed, rows=1):
    return torch.from_numpy(synthetic.complex_uniform(seed, (rows, n))).to("cuda")


def close(a, b, tol=TOL):
    torch.cuda.synchronize()
    if a.numel() == 0 and b.numel() == 0:
        return a.shape == b.shape
    return a.shape == b.shape and float((a - b).abs().max()) <= tol


def rev(x):
    """Cyclic.reverse (Cyclic.hs:20-28): x[(-k) mod n]."""
    n = x.shape[-1]
    if n == 0:
        return x
    return torch.roll(torch.flip(x, [-1]), 1, -1)


@pytest.mark.parametrize("n", LENS)
@pytest.mark.parametrize("s", [-1, 1])
def test_homogeneity_additivity(n, s):
    T = planned.Transform(Sign(s), n)
    x, y = c(n, 1), c(n, 2)
    a = complex(0.3, -0.8)
    assert close(T(a * x), a * T(x))                          # Fourier.hs:355-360
    assert close(T(x) + T(y), T(x + y))                       # Fourier.hs:361-366


@pytest.mark.parametrize("n", [n for n in LENS if n > 0])
def test_basis_vector(n):
    """Fourier.hs:367-376: transform inverse of unit vector k is cis(2 pi k j / n)."""
    for k in sorted({0, 1 % n, n // 2, n - 1}):
        e = torch.zeros(1, n, dtype=torch.complex128, device="cuda")
        e[0, k] = 1
        j = np.arange(n)
        expect = torch.from_numpy(np.exp(2j * np.pi * ((k * j) % n) / n)[None]).to("cuda")
        assert close(planned.transform(inverse, n)(e), expect)


@pytest.mark.parametrize("n", LENS)
def test_inverse_double_reverse_conjugate(n):
    x = c(n, 3)
    assert close(x, scale_down(planned.transform(forward, n)(planned.transform(inverse, n)(x))))  # 377-382
    for s in (-1, 1):
        T = planned.Transform(Sign(s), n)
        assert close(x, rev(scale_down(T(T(x)))))             # 383-390
        assert close(rev(T(x)), T(rev(x)))                     # 391-395
        assert close(torch.conj(T(x)), T(torch.conj(rev(x)).contiguous()))  # 396-402


@pytest.mark.parametrize("n", LENS)
def test_isometry_unitarity(n):
    x, y = c(n, 4), c(n, 5)
    for s in (-1, 1):
        T = planned.Transform(Sign(s), n)
        assert abs(float(T(x).norm()) - np.sqrt(n) * float(x.norm())) <= TOL   # 403-409
        lhs = (T(x) * torch.conj(T(y))).sum()
        rhs = n * (x * torch.conj(y)).sum()
        assert abs(complex(lhs) - complex(rhs)) <= TOL                          # 410-417


@pytest.mark.parametrize("n", [n for n in LENS if n > 0])
def test_convolution_laws(n):
    x, y, z = c(n, 6), c(n, 7), c(n, 8)
    conv = planned.convolve_cyclic(n)
    assert close(conv(x, y), conv(y, x))                                  # 418-423
    assert close(conv(conv(x, y), z), conv(x, conv(y, z)))                # 424-429
    assert close(conv(x, y + z), conv(x, y) + conv(x, z))                 # 430-435


@pytest.mark.parametrize("n", [n for n in LENS if n > 0])
def test_karatsuba_vs_fft_convolution(n):
    """Fourier.hs:447-462 with integer-valued complex operands, so the Int64
    Karatsuba (Gauss' three products, Adhoc.hs:114-133) is exact and the FP64
    FFT convolution must agree to 1e-10."""
    r = np.random.default_rng(n)
    xr, xi, yr, yi = (r.integers(-8, 8, (1, n)) for _ in range(4))
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    k = convolution.karatsuba
    a, b, m = k(t(xr), t(yr)), k(t(xi), t(yi)), k(t(xr + xi), t(yr + yi))
    lin = torch.complex((a - b).double(), (m - (a + b)).double())
    L = 2 * n - 1
    pad = lambda u: torch.nn.functional.pad(u, (0, L - n))
    xc = torch.complex(t(xr).double(), t(xi).double())
    yc = torch.complex(t(yr).double(), t(yi).double())
    assert close(lin, planned.convolve_cyclic(L)(pad(xc).contiguous(), pad(yc).contiguous()))
    cyc = lambda p, q: convolution.cyclic(k)(t(p), t(q))
    cyc_g = torch.complex((cyc(xr, yr) - cyc(xi, yi)).double(),
                          (cyc(xr + xi, yr + yi) - cyc(xr, yr) - cyc(xi, yi)).double())
    assert close(cyc_g, planned.convolve_cyclic(n)(xc, yc))


@pytest.mark.parametrize("n", [n for n in LENS if n > 0 and n % 2 == 0] + [2, 128, 200])
def test_real_spectrum_properties(n):
    """Fourier.hs:463-496."""
    x = c(n, 9)
    xr = x.real.contiguous()
    cr = lambda v: torch.complex(v, torch.zeros_like(v))
    h = n // 2
    assert close(real.to_spectrum(planned.transform(forward, h), xr), planned.transform(forward, n)(cr(xr)))
    sa = x + torch.conj(rev(x))                                              # toSelfAdjoint
    assert close(cr(real.from_spectrum(planned.transform(inverse, h), sa)), planned.transform(inverse, n)(sa))
    back = real.from_spectrum(planned.transform(inverse, h), real.to_spectrum(planned.transform(forward, h), xr))
    assert close(back / n, xr)
    fwd = real.to_spectrum(planned.transform(forward, h), real.from_spectrum(planned.transform(inverse, h), sa))
    assert close(fwd / n, sa)


@pytest.mark.parametrize("n", LENS)
def test_generic_vs_chirp_and_decompose(n):
    """Fourier.hs:345-354: Planned.transform vs transformChirp2 / transformChirp235 /
    transformDecompose, at the reference's 1e-10."""
    x = c(n, 10)
    for s in (-1, 1):
        g = planned.Transform(Sign(s), n)(x)
        assert close(g, planned.transform_chirp2(Sign(s), n)(x))
        assert close(g, planned.transform_chirp235(Sign(s), n)(x))
        assert close(g, planned.transform_decompose(Sign(s), n)(x))


def test_random_lengths_against_numpy():
    """Seeded random lengths over every plan kind the planner picks (single
    pass, prime-factor, four-step, Rader, Bluestein, nested sizes up to 4e5),
    both precisions and signs, against numpy's FP64 FFT of the same inputs:
    relative L2 <= 2e-6 in FP32 (inputs rounded to FP32 first) and 1e-12 in
    FP64."""
    rng = np.random.default_rng(7)
    lens = sorted(set(int(n) for n in np.concatenate([rng.integers(1, 40000, 120), rng.integers(40000, 400000, 12)])))
    bad = []
    for i, n in enumerate(lens):
        rows = 2 if n < 100000 else 1
        x64 = synthetic.complex_uniform(100 + i, (rows, n))
        for dt, tol in ((torch.complex64, 2e-6), (torch.complex128, 1e-12)):
            xt = torch.from_numpy(x64).to("cuda", dt)
            xh = xt.cpu().numpy().astype(np.complex128)
            for sgn in (forward, inverse):
                ref = np.fft.fft(xh, axis=-1) if sgn.value < 0 else np.fft.ifft(xh, axis=-1) * n
                got = planned.transform(sgn, n)(xt).cpu().numpy().astype(np.complex128)
                err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
                if not err <= tol:
                    bad.append((n, str(dt), sgn.value, float(err), planned.describe(n)))
    print(f"{len(lens)} lengths from {lens[0]} to {lens[-1]}, both precisions and signs: {len(bad)} above tolerance")
    assert not bad, bad[:5]
