"""The Python mirror validates every operand and caller-supplied result before
it reaches a kernel: a tensor of the wrong shape, dtype or size would make a
kernel that indexes by the call's extents read or write outside it.  Runs on
the CPU: the device check is relaxed to accept host tensors and the C library
is replaced by a stub that records calls, so a test sees whether a call got
through validation without launching anything.
"""
import pytest

torch = pytest.importorskip("torch")

from accelerate_fourier import _abi, _arrays, batch, convolution, planned, real  # noqa: E402
from accelerate_fourier.sign import forward  # noqa: E402


class _StubLib:
    """Every af_* call succeeds and is recorded; nothing runs."""

    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        def call(*args):
            self.calls.append(name)
            return 0
        return call


@pytest.fixture
def stub(monkeypatch):
    lib = _StubLib()
    monkeypatch.setattr(_abi, "lib", lambda: lib)

    def host_ok(t, what):
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{what} must be a torch tensor")
        if not t.is_contiguous():
            raise ValueError(f"{what} must be contiguous")
    monkeypatch.setattr(_arrays, "_check_tensor", host_ok)
    monkeypatch.setattr(_arrays, "stream_handle", lambda: 0)
    planned._plan_cache.clear()
    yield lib
    planned._plan_cache.clear()


def test_transform_out_checked(stub):
    t = planned.transform(forward, 8)
    x = torch.zeros(3, 8, dtype=torch.complex64)
    t(x, out=torch.empty(3, 8, dtype=torch.complex64))
    assert stub.calls[-1] == "af_execute"
    n = len(stub.calls)
    with pytest.raises(ValueError):
        t(x, out=torch.empty(2, 8, dtype=torch.complex64))
    with pytest.raises(TypeError):
        t(x, out=torch.empty(3, 8, dtype=torch.complex128))
    with pytest.raises(ValueError):
        t(x, out=(torch.empty(3, 8), torch.empty(3, 7)))
    with pytest.raises(ValueError):
        t(x, out=torch.empty(3, 8, dtype=torch.complex64).conj())
    assert len(stub.calls) == n
    t((torch.zeros(3, 8), torch.zeros(3, 8)), out=(torch.empty(3, 8), torch.empty(3, 8)))
    assert stub.calls[-1] == "af_execute"


def test_convolve_and_vjp_operands_checked(stub):
    x = torch.zeros(4, 16)
    real.convolve(x, torch.zeros(4, 16))
    real.convolve_vjp(x, torch.zeros(4, 16), torch.zeros(4, 16))
    n = len(stub.calls)
    with pytest.raises(ValueError):
        real.convolve(x, torch.zeros(4, 16), out=torch.empty(4, 8))
    with pytest.raises(ValueError):  # kernel shorter than the signal
        real.convolve_vjp(x, torch.zeros(4, 8), torch.zeros(4, 16))
    with pytest.raises(ValueError):  # cotangent with fewer rows
        real.convolve_vjp(x, torch.zeros(4, 16), torch.zeros(2, 16))
    with pytest.raises(ValueError):  # cotangent of another precision
        real.convolve_vjp(x, torch.zeros(4, 16), torch.zeros(4, 16, dtype=torch.float64))
    with pytest.raises(ValueError):
        real.convolve_vjp(x, torch.zeros(4, 16), torch.zeros(4, 16), dx=torch.empty(4, 15))
    with pytest.raises(TypeError):
        real.convolve_vjp(x, torch.zeros(4, 16), torch.zeros(4, 16), dk=torch.empty(4, 16, dtype=torch.float64))
    assert len(stub.calls) == n


def test_real_spectrum_outs_checked(stub):
    x = torch.zeros(2, 16)
    sub = planned.transform(forward, 8)
    real.to_spectrum(sub, x)
    n = len(stub.calls)
    with pytest.raises(ValueError):
        real.to_spectrum(sub, x, out=torch.empty(2, 8, dtype=torch.complex64))
    with pytest.raises(ValueError):
        real.from_spectrum(sub, torch.zeros(2, 16, dtype=torch.complex64), out=torch.empty(2, 15))
    with pytest.raises(ValueError):
        real.two_to_spectrum(planned.transform(forward, 16), x, x,
                             out=(torch.empty(2, 16, dtype=torch.complex64), torch.empty(1, 16, dtype=torch.complex64)))
    assert len(stub.calls) == n


def test_karatsuba_sizes_checked(stub):
    x = torch.zeros(3, 10, dtype=torch.int64)
    convolution.karatsuba(x, x)
    n = len(stub.calls)
    with pytest.raises(ValueError):
        convolution.karatsuba(x, x, out=torch.empty(3, 18, dtype=torch.int64))
    with pytest.raises(ValueError):
        convolution.cyclic(convolution.karatsuba)(x, x, out=torch.empty(3, 9, dtype=torch.int64))
    flat = torch.zeros(20, dtype=torch.int64)
    convolution.karatsuba_ragged([0, 5, 20], flat, flat)
    n = len(stub.calls)
    with pytest.raises(ValueError):  # offsets run past the operands
        convolution.karatsuba_ragged([0, 5, 21], flat, flat)
    with pytest.raises(ValueError):  # an empty pair
        convolution.karatsuba_ragged([0, 5, 5, 20], flat, flat)
    with pytest.raises(ValueError):  # result too short: 2 * 20 - 2 = 38 words needed
        convolution.karatsuba_ragged([0, 5, 20], flat, flat, out=torch.empty(37, dtype=torch.int64))
    assert len(stub.calls) == n


def test_uneven_batch_checks_out_and_resolves_conjugate_views(stub):
    b = batch.UnevenBatch([4, 7], forward, "f32")
    x = torch.zeros(11, dtype=torch.complex64)
    n = len(stub.calls)
    with pytest.raises(ValueError):
        b(x, out=torch.empty(10, dtype=torch.complex64))
    assert len(stub.calls) == n
    seen = {}
    orig = _arrays.complex_view

    def spy(t, what="input"):
        seen["conj"] = bool(getattr(t, "is_conj", lambda: False)())
        return orig(t, what)
    _arrays.complex_view = spy
    try:
        b(torch.ones(11, dtype=torch.complex64).conj())
    finally:
        _arrays.complex_view = orig
    assert seen["conj"] is False and stub.calls[-1] == "af_batch_execute"
