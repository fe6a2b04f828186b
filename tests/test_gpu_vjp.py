"""Adjoints through the C ABI: gradcheck (central finite differences,
h = 1e-6, FP64), the closed forms (conjugate-transposed DFT, correlation) and
the reference's inverse-transform gradient computed by the oracle, at the real
convolution's full configuration (512 x 8192, seed 3)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from accelerate_fourier import autograd as AG, planned, real, synthetic  # noqa: E402
from accelerate_fourier.sign import Sign, forward  # noqa: E402


def dev(x, dt=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda", dt)


@pytest.mark.parametrize("n", [1, 2, 8, 12, 17, 47, 64, 97])
@pytest.mark.parametrize("sign", [-1, 1])
def test_transform_gradcheck(n, sign):
    x = dev(synthetic.complex_uniform(n, (2, n)), torch.complex128).requires_grad_()
    t = planned.Transform(Sign(sign), n)
    assert torch.autograd.gradcheck(lambda v: AG.TransformFn.apply(v, t), (x,), eps=1e-6, atol=1e-8)


@pytest.mark.parametrize("n", [2, 4, 16, 30, 64])
def test_real_packing_gradcheck(n):
    x = dev(synthetic.normal(n, 2 * n).reshape(2, n)).requires_grad_()
    assert torch.autograd.gradcheck(AG.ToSpectrumFn.apply, (x,), eps=1e-6, atol=1e-8)
    s = dev(synthetic.complex_uniform(n + 1, (2, n)), torch.complex128).requires_grad_()
    assert torch.autograd.gradcheck(AG.FromSpectrumFn.apply, (s,), eps=1e-6, atol=1e-8)


@pytest.mark.parametrize("n", [2, 8, 64, 256])
def test_real_conv_gradcheck(n):
    x = dev(synthetic.normal(1, 2 * n).reshape(2, n)).requires_grad_()
    k = dev(synthetic.normal(2, 2 * n).reshape(2, n) / math.sqrt(n)).requires_grad_()
    assert torch.autograd.gradcheck(AG.RealConvFn.apply, (x, k), eps=1e-6, atol=1e-8)


def test_real_conv_adjoint_config():
    """512 x 8192, x ~ N(0,1), k ~ N(0, 1/8192), w ~ N(0,1), seed 3: the fused
    adjoint against the inverse-transform gradient (oracle, FP64) on sampled
    rows, random-direction central differences (h = 1e-6) on the full batch,
    and the FP32 adjoint against the FP64 one."""
    import oracle
    B, N = 512, 8192
    allv = synthetic.normal(3, 3 * B * N)
    xs = allv[:B * N].reshape(B, N)
    ks = allv[B * N:2 * B * N].reshape(B, N) / math.sqrt(N)
    ws = allv[2 * B * N:].reshape(B, N)
    x, k, w = dev(xs), dev(ks), dev(ws)
    dx, dk = real.convolve_vjp(x, k, w)
    torch.cuda.synchronize()
    rows = [0, 17, 511]
    Sx, Sk, Sw = (oracle.to_spectrum(a[rows]) for a in (xs, ks, ws))
    ref_dx = oracle.from_spectrum(Sw * np.conj(Sk))
    ref_dk = oracle.from_spectrum(Sw * np.conj(Sx))
    got_dx, got_dk = dx.cpu().numpy()[rows], dk.cpu().numpy()[rows]
    assert np.linalg.norm(got_dx - ref_dx) / np.linalg.norm(ref_dx) < 1e-12
    assert np.linalg.norm(got_dk - ref_dk) / np.linalg.norm(ref_dk) < 1e-12
    # directional derivative by central differences on the whole batch
    d = dev(synthetic.normal(99, B * N).reshape(B, N))
    h = 1e-6
    loss = lambda xx, kk: float((w * real.convolve(xx, kk)).sum())
    fd_x = (loss(x + h * d, k) - loss(x - h * d, k)) / (2 * h)
    fd_k = (loss(x, k + h * d) - loss(x, k - h * d)) / (2 * h)
    an_x, an_k = float((dx * d).sum()), float((dk * d).sum())
    assert abs(fd_x - an_x) <= 1e-6 * abs(an_x)
    assert abs(fd_k - an_k) <= 1e-6 * abs(an_k)
    # FP32 adjoint against the FP64 one
    dx32, dk32 = real.convolve_vjp(x.float(), k.float(), w.float())
    torch.cuda.synchronize()
    e = (dx32.double() - dx).norm() / dx.norm()
    assert float(e) < 1e-6


def test_transform_vjp_is_opposite_sign():
    for n in (1024, 65536, 65521, 100003):
        g = dev(synthetic.complex_uniform(n, (2, n)), torch.complex128)
        t = planned.transform(forward, n)
        a = t.vjp(g)
        b = planned.transform(Sign(1), n)(g)
        torch.cuda.synchronize()
        assert float((a - b).norm() / b.norm()) < 1e-13
