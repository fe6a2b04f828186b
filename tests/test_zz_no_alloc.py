"""af_execute_noalloc through the C ABI: the allocation-free variant of
af_execute (scratch from the library's fixed per-GPU buffer) gives the same
bits as af_execute with its own workspace, on two calls in a row per plan, the
second with fewer rows, so the second call runs over scratch the first one
left behind.  Plans with scratch: the 64 x 65536 four-step (32 MiB of scratch
in FP32), a Bluestein prime and a Rader prime with four-step cores, both
precisions and layouts; a call beyond the buffer is refused before any launch.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from accelerate_fourier import _abi, planned, synthetic  # noqa: E402


def _run(plan, batch, x, planar, noalloc):
    """One call on the current stream; returns the result on the host."""
    lib = _abi.lib()
    s = torch.cuda.current_stream().cuda_stream
    if planar:
        xin = (x.real.contiguous(), x.imag.contiguous())
        out = (torch.empty_like(xin[0]), torch.empty_like(xin[1]))
        a = (xin[0].data_ptr(), xin[1].data_ptr(), out[0].data_ptr(), out[1].data_ptr())
    else:
        xin = x
        out = torch.empty_like(x)
        a = (x.data_ptr(), None, out.data_ptr(), None)
    if noalloc:
        st = lib.af_execute_noalloc(plan.handle, batch, a[0], a[1], a[2], a[3], s)
    else:
        ws = torch.empty(max(plan.workspace_size(batch), 1), dtype=torch.uint8, device="cuda")
        st = lib.af_execute(plan.handle, batch, a[0], a[1], a[2], a[3], ws.data_ptr(), s)
    _abi.check(st, "af_execute_noalloc" if noalloc else "af_execute")
    torch.cuda.synchronize()
    if planar:
        return out[0].cpu().numpy() + 1j * out[1].cpu().numpy()
    return out.cpu().numpy()


@pytest.mark.parametrize("n,prec,rows", [(65536, _abi.F32, (64, 8)), (65536, _abi.F64, (16, 3)),
                                         (100003, _abi.F32, (4, 1)), (65521, _abi.F64, (3, 2))])
@pytest.mark.parametrize("planar", [True, False])
def test_noalloc_matches_execute(n, prec, rows, planar):
    plan = planned.get_plan(n, -1, _abi.C2C, prec)
    assert plan.workspace_size(rows[0]) > 0
    ct = torch.complex64 if prec == _abi.F32 else torch.complex128
    for k, batch in enumerate(rows):
        x = torch.from_numpy(synthetic.complex_normal(40 + k + n % 97, (batch, n))).to("cuda", ct)
        ref = _run(plan, batch, x, planar, noalloc=False)
        got = _run(plan, batch, x, planar, noalloc=True)
        assert np.array_equal(got, ref), (n, prec, batch, planar)
        if n == 65536 and k == 0:
            assert np.linalg.norm(got[0] - np.fft.fft(x[0].cpu().numpy())) <= 1e-6 * np.linalg.norm(got[0])


def test_noalloc_refuses_beyond_buffer():
    """One 2^24 FP32 transform needs 128 MiB of scratch, over the 64 MiB buffer:
    AF_UNSUPPORTED with the sizes in af_last_error, and the output untouched."""
    plan = planned.get_plan(1 << 24, -1, _abi.C2C, _abi.F32)
    assert plan.workspace_size(1) > 64 << 20
    x = torch.zeros(1 << 24, dtype=torch.complex64, device="cuda")
    y = torch.full_like(x, 7)
    st = _abi.lib().af_execute_noalloc(plan.handle, 1, x.data_ptr(), None, y.data_ptr(), None,
                                       torch.cuda.current_stream().cuda_stream)
    assert st == _abi.UNSUPPORTED
    assert b"fixed buffer" in _abi.lib().af_last_error()
    torch.cuda.synchronize()
    assert bool((y == 7).all())
