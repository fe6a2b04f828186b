"""Benchmark of the accelerate-fourier transform path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Headline workload: batched power-of-two complex FFT in Complex Float, 64 rows
of length 65536 (FP32 storage and arithmetic, FP64-generated hi/lo twiddles),
forward, planar (Accelerate's SoA) layout, inputs N(0,1) from seed 1, resident
in HBM.  One step = one forward transform of the whole batch.  Prints one JSON
line: GFLOP/s at 5 N log2 N, the roofline of the dominant kernel (HIP events on
the stream the kernels run on), the CPU baseline (the oracle, a C restatement
of the reference's algorithm, on host cores), and side measurements (plain
FP64 path, 2^24, the FP64 256 x 1024 case, the real convolution with its
adjoint, Karatsuba, and the error of the FP32 path against FP64).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import re
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(REPO, "accelerate-fourier_amd"), os.path.join(REPO, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
# the two launches of the 64 x 65536 four-step transform (65536 = 256 x 256)
PASS_NAMES = {0: "pass 0: 256-point column FFTs + four-step twiddles",
              1: "pass 1: 256-point row FFTs, transposed store"}
METRIC = "FP32 batched FFT GFLOP/s(5NlgN) & %HBM peak vs FP64 plain; fwd+bwd conv/s"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--skip-extras", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="run the headline transform with the planner's tile shapes instead of autotuning "
                         "them first (af_plan_tune)")
    ap.add_argument("--profile", action="store_true",
                    help="only the headline transform (steps launches) and a 32 MiB device copy of "
                         "known size, for rocprofv3 kernel traces and counters")
    ap.add_argument("--profile-uneven", action="store_true",
                    help="only the 4096-row uneven batch (FP32 forward, chunk size autotuned unless "
                         "--no-tune, warmup + steps runs), for rocprofv3 kernel traces")
    ap.add_argument("--profile-conv", action="store_true",
                    help="only the 512 x 8192 FP32 real convolution and its adjoint (warmup + steps "
                         "calls), for rocprofv3 kernel traces")
    ap.add_argument("--profile-c2b", action="store_true",
                    help="only one 2^24 FP32 transform, autotuned unless --no-tune (warmup + steps runs), "
                         "and a 128 MiB device copy of known size, for rocprofv3 kernel traces and counters")
    ap.add_argument("--profile-c1", action="store_true",
                    help="only the FP64 256 x 1024 transform, autotuned (warmup + steps runs), for rocprofv3 "
                         "kernel traces")
    ap.add_argument("--profile-karatsuba", action="store_true",
                    help="only the 1024-pair ragged Int64 Karatsuba (warmup + steps calls), for "
                         "rocprofv3 kernel traces")
    return ap.parse_args()


class C2C:
    """A C2C plan called straight through the ABI with a preallocated workspace."""

    def __init__(self, n, batch, prec, sign=-1, planar=True):
        import torch
        from accelerate_fourier import _abi, planned

        self.torch, self.abi = torch, _abi
        self.n, self.batch, self.prec, self.planar = n, batch, prec, planar
        self.plan = planned.Plan(n, sign, _abi.C2C, prec)
        ws = self.plan.workspace_size(batch)
        self.ws = torch.empty(max(ws, 16), dtype=torch.uint8, device="cuda")
        cnt = ctypes.c_int32()
        _abi.check(_abi.lib().af_plan_num_passes(self.plan.handle, batch, ctypes.byref(cnt)), "num_passes")
        self.npasses = cnt.value

    def set_io(self, x):
        """x: complex numpy array (batch, n)."""
        torch = self.torch
        rt = torch.float64 if self.prec == self.abi.F64 else torch.float32
        ct = torch.complex128 if self.prec == self.abi.F64 else torch.complex64
        if self.planar:
            self.xin = (torch.from_numpy(x.real.copy()).to("cuda", rt), torch.from_numpy(x.imag.copy()).to("cuda", rt))
            self.xout = (torch.empty_like(self.xin[0]), torch.empty_like(self.xin[1]))
            self.ptrs = (self.xin[0].data_ptr(), self.xin[1].data_ptr(), self.xout[0].data_ptr(), self.xout[1].data_ptr())
        else:
            self.xin = torch.from_numpy(x).to("cuda", ct)
            self.xout = torch.empty_like(self.xin)
            self.ptrs = (self.xin.data_ptr(), None, self.xout.data_ptr(), None)

    def run(self):
        s = self.torch.cuda.current_stream().cuda_stream
        a = self.ptrs
        st = self.abi.lib().af_execute(self.plan.handle, self.batch, a[0], a[1], a[2], a[3], self.ws.data_ptr(), s)
        self.abi.check(st, "af_execute")

    def tune(self):
        """af_plan_tune for this batch and layout; returns the tuner's report."""
        buf = ctypes.create_string_buffer(8192)
        s = self.torch.cuda.current_stream().cuda_stream
        self.abi.check(self.abi.lib().af_plan_tune(self.plan.handle, self.batch, 1 if self.planar else 0, s, buf,
                                                   len(buf)), "af_plan_tune")
        return buf.value.decode()

    def run_pass(self, i):
        s = self.torch.cuda.current_stream().cuda_stream
        a = self.ptrs
        st = self.abi.lib().af_execute_pass(self.plan.handle, i, self.batch, a[0], a[1], a[2], a[3],
                                            self.ws.data_ptr(), s)
        self.abi.check(st, "af_execute_pass")

    def result(self):
        self.torch.cuda.synchronize()
        if self.planar:
            return self.xout[0].cpu().numpy() + 1j * self.xout[1].cpu().numpy()
        return self.xout.cpu().numpy()


def tile_lines(report, i, default):
    """Lines per tile the autotuner chose for launch i (its report line), else default."""
    m = re.search(rf"^pass {i} L=\d+: B=(\d+)", report, re.M)
    return int(m.group(1)) if m else default


def launch_name(report, i):
    return f"k_fast<float, 256, {tile_lines(report, i, 16)}, -1> " + PASS_NAMES.get(i, f"pass {i}")


def time_it(fn, steps, warmup):
    """Average ms per call, HIP events on the current stream around `steps` calls."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / steps


def profiled_traffic(pass_index):
    """HBM bytes per launch of one pass of the headline transform, from the
    committed rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE, scaled by
    what the same counters read for a copy of known size); None if absent."""
    path = os.path.join(REPO, "profiles", "c2a_64x65536_f32_hbm_counters.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(f"pass {pass_index} traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def karatsuba_subtree_macs(L):
    """Multiply-adds of the matrix-core Karatsuba kernel's subtree products for
    one pair of length L: padded to P = 2^d s (s <= 64, d smallest), T = d - 3
    top Karatsuba levels (none below d = 3) and 3^T schoolbook products of
    Q x Q, Q = 2^min(d, 3) s."""
    d, s = 0, L
    while s > 64:
        d += 1
        s = (L + (1 << d) - 1) >> d
    B = min(d, 3)
    q = s << B
    return (3 ** (d - B)) * q * q


def profiled_c2b_traffic():
    """HBM bytes of one 2^24 transform from the committed counter summary, or None."""
    path = os.path.join(REPO, "profiles", "c2b_2p24_f32_hbm_counters.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("transform traffic_bytes:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def conv_workload(torch):
    """Real cyclic convolution and its adjoint, FP32 512 rows x 8192: x, the
    kernel k (scaled by 1/sqrt(N)) and the cotangent w from normal draws of
    splitmix64 seeds 3, 4, 5, resident in HBM.  One call is the forward
    convolution and the vjp for both operands."""
    from accelerate_fourier import real, synthetic
    N = 8192
    xr = torch.from_numpy(synthetic.normal(3, 512 * N).reshape(512, N)).to("cuda", torch.float32)
    kr = torch.from_numpy(synthetic.normal(4, 512 * N).reshape(512, N) / math.sqrt(N)).to("cuda", torch.float32)
    wr = torch.from_numpy(synthetic.normal(5, 512 * N).reshape(512, N)).to("cuda", torch.float32)
    yr, dx, dk = torch.empty_like(xr), torch.empty_like(xr), torch.empty_like(xr)

    def fb():
        real.convolve(xr, kr, out=yr)
        real.convolve_vjp(xr, kr, wr, dx, dk)
    return fb, N


def karatsuba_workload(torch, np):
    """Int64 Karatsuba over 1024 ragged pairs, lengths U[256, 4096] (seed 4),
    entries U[-2^15, 2^15) (seeds 40, 41), resident in HBM.  Returns the call
    and the lengths."""
    from accelerate_fourier import convolution, synthetic
    lens = synthetic.integers(4, 1024, 256, 4097)
    off = np.concatenate([[0], np.cumsum(lens)])
    xa = torch.from_numpy(synthetic.integers(40, int(off[-1]), -2 ** 15, 2 ** 15)).to("cuda")
    ya = torch.from_numpy(synthetic.integers(41, int(off[-1]), -2 ** 15, 2 ** 15)).to("cuda")
    za = torch.empty(2 * int(off[-1]) - 1024, dtype=torch.int64, device="cuda")
    return (lambda: convolution.karatsuba_ragged(off, xa, ya, out=za)), lens, (off, xa, ya, za)


def karatsuba_fft_f32(torch, np, lens, data, steps, warmup):
    """The Karatsuba pairs cast to FP32 and convolved by the fused real FFT
    convolution: pair p is zero-padded to M = the next power of two >= 2L - 1
    (one launch per M) with x scaled by 1/M (exact), so the result is the
    linear convolution.  Against the exact Int64 products (the Karatsuba
    result of the same call): time, max-abs error, and the fraction of
    coefficients that rounding the FP32 result to the nearest integer recovers
    (values reach 2^36 and more, far past FP32's 24-bit significand)."""
    from accelerate_fourier import real
    off, xa, ya, za = data
    xh, yh = xa.cpu().numpy(), ya.cpu().numpy()
    Ms = np.array([1 << int(np.ceil(np.log2(2 * L - 1))) for L in lens])
    groups = []
    for M in sorted(set(Ms.tolist())):
        idx = np.nonzero(Ms == M)[0]
        xp = np.zeros((len(idx), M), np.float32)
        yp = np.zeros((len(idx), M), np.float32)
        for j, p in enumerate(idx):
            L = int(lens[p])
            xp[j, :L] = xh[off[p]:off[p] + L].astype(np.float32) / M
            yp[j, :L] = yh[off[p]:off[p] + L].astype(np.float32)
        xd, yd = torch.from_numpy(xp).cuda(), torch.from_numpy(yp).cuda()
        groups.append((idx, xd, yd, torch.empty_like(xd)))

    def run():
        for _, xd, yd, zd in groups:
            real.convolve(xd, yd, out=zd)
    ms = time_it(run, steps, warmup)
    run()
    zx = za.cpu().numpy()
    err = big = 0.0
    hit = tot = 0
    for idx, _, _, zd in groups:
        zf = zd.cpu().numpy()
        for j, p in enumerate(idx):
            L = int(lens[p])
            got = zf[j, :2 * L - 1].astype(np.float64)
            ex = zx[2 * off[p] - p:2 * off[p] - p + 2 * L - 1]
            err = max(err, float(np.max(np.abs(got - ex))))
            big = max(big, float(np.max(np.abs(ex))))
            hit += int(np.sum(np.rint(got).astype(np.int64) == ex))
            tot += 2 * L - 1
    return {"c5_fft_f32_ms": ms, "c5_fft_f32_launches": len(groups), "c5_fft_f32_max_abs_err": err,
            "c5_exact_max_abs": big, "c5_fft_f32_max_abs_err_rel_to_max": err / big,
            "c5_fft_f32_rounded_exact_frac": hit / tot}


def uneven_batch(torch, np, steps, tune=True):
    """The 4096-row uneven batch (four length classes, sum of lengths 2.86e8), FP32
    forward, inputs U[-1, 1) from splitmix64 seed 2 resident in HBM, its graph's
    chunk size autotuned (af_batch_tune) unless tune is False; error of every row
    against the FP64 batch on the same inputs, summarised per class."""
    from accelerate_fourier import batch, forward, synthetic
    lens, cls = synthetic.uneven_lengths(2)
    tot = int(lens.sum())
    t0 = time.perf_counter()
    b32 = batch.UnevenBatch(lens, forward, "f32")
    b64 = batch.UnevenBatch(lens, forward, "f64")
    create_s = time.perf_counter() - t0
    default_plan = b32.describe().split("\n")[0]
    report = b32.tune() if tune else ""
    v = synthetic.uniform(2, 2 * tot)
    x64 = torch.complex(torch.from_numpy(v[:tot]), torch.from_numpy(v[tot:])).cuda()
    del v
    x32 = x64.to(torch.complex64)
    y32 = b32(x32)
    ms = time_it(lambda: b32(x32, out=y32), max(steps // 4, 3), 1)
    y64 = b64(x64)
    torch.cuda.synchronize()
    # per-row norms over views of each row (scatter-adds into 4096 bins take a
    # minute in FP64 atomics)
    err = (y32.to(torch.complex128) - y64).abs()
    sizes = [int(n) for n in lens]
    d = torch.stack([torch.linalg.vector_norm(r) for r in torch.split(err, sizes)])
    nrm = torch.stack([torch.linalg.vector_norm(r) for r in torch.split(y64, sizes)])
    mx = torch.stack([r.amax() for r in torch.split(err, sizes)]).cpu().numpy()
    rel = (d / nrm).cpu().numpy()
    del err
    flops = 5.0 * float(np.sum(lens * np.log2(np.maximum(lens, 2))))
    out = {"c3_uneven_rows": int(len(lens)), "c3_uneven_total_length": tot, "c3_uneven_f32_ms": ms,
           "c3_uneven_GFLOPs": flops / (ms * 1e-3) / 1e9,
           "c3_uneven_hbm_frac": 16.0 * tot / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "c3_uneven_create_s": create_s, "c3_uneven_plan": b32.describe().split("\n")[0],
           "c3_uneven_default_plan": default_plan,
           "c3_autotune": report.strip().split("\n") if report else None, "c3_uneven_error_vs_fp64": {}}
    for c, name in enumerate(synthetic.UNEVEN_CLASSES):
        r = rel[cls == c]
        out["c3_uneven_error_vs_fp64"][name] = {"rows": int(len(r)), "max_rel_l2": float(r.max()),
                                                 "median_rel_l2": float(np.median(r)),
                                                 "max_abs": float(mx[cls == c].max())}
    del b32, b64, x32, x64, y32, y64
    return out


def rel_l2(a, b):
    import numpy as np
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: the transform path shards across independent rows "
                                   "without communication, so its figure is the one-GPU throughput",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch
    from accelerate_fourier import _abi, synthetic

    assert torch.cuda.is_available(), "bench.py needs the GPU"
    torch.cuda.init()
    steps, warmup = args.steps, args.warmup
    if args.profile_uneven:
        from accelerate_fourier import batch as ubatch, forward, synthetic as syn
        lens, _ = syn.uneven_lengths(2)
        tot = int(lens.sum())
        b32 = ubatch.UnevenBatch(lens, forward, "f32")
        report = "" if args.no_tune else b32.tune()
        v = syn.uniform(2, 2 * tot)
        x32 = torch.complex(torch.from_numpy(v[:tot].astype(np.float32)),
                            torch.from_numpy(v[tot:].astype(np.float32))).cuda()
        del v
        y = b32(x32)
        for _ in range(warmup + steps):
            b32(x32, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"profile": "done", "uneven_runs": 1 + warmup + steps, "plan": b32.describe(),
                          "autotune": report.strip().split("\n") if report else None}))
        return
    if args.profile_conv or args.profile_karatsuba:
        fn = conv_workload(torch)[0] if args.profile_conv else karatsuba_workload(torch, np)[0]
        for _ in range(warmup + steps):
            fn()
        torch.cuda.synchronize()
        print(json.dumps({"profile": "done", "calls": warmup + steps,
                          "workload": "conv" if args.profile_conv else "karatsuba"}))
        return
    if args.profile_c2b or args.profile_c1:
        if args.profile_c2b:
            g = C2C(1 << 24, 1, _abi.F32)
            g.set_io(synthetic.complex_normal(1, (1, 1 << 24)))
        else:
            g = C2C(1024, 256, _abi.F64)
            g.set_io(synthetic.complex_uniform(0, (256, 1024)))
        report = "" if args.no_tune else g.tune()
        for _ in range(warmup + steps):
            g.run()
        if args.profile_c2b:
            src = torch.empty(128 * 2 ** 20 // 4, dtype=torch.float32, device="cuda")
            dst = torch.empty_like(src)
            for _ in range(steps):
                dst.copy_(src)  # reference kernel: reads and writes exactly 128 MiB
        torch.cuda.synchronize()
        print(json.dumps({"profile": "done", "workload": "c2b 2^24 f32" if args.profile_c2b else "c1 256x1024 f64",
                          "transform_launches": g.npasses * (warmup + steps),
                          "autotune": report.strip().split("\n") if report else None}))
        return
    n, batch = 65536, 64
    flops = 5.0 * n * math.log2(n) * batch
    alg_bytes = 2 * 8 * n * batch  # one read and one write of the C32 array

    x = synthetic.complex_normal(1, (batch, n))
    f32 = C2C(n, batch, _abi.F32)
    f32.set_io(x)
    tune_report = "" if args.no_tune else f32.tune()
    if args.profile:
        for _ in range(warmup + steps):
            f32.run()
        src = torch.empty(32 * 2 ** 20 // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(steps):
            dst.copy_(src)  # reference kernel: reads and writes exactly 32 MiB
        torch.cuda.synchronize()
        print(json.dumps({"profile": "done", "transform_launches": f32.npasses * (warmup + steps),
                          "copy_bytes_each_way": 32 * 2 ** 20}))
        return
    ms = time_it(f32.run, steps, warmup)
    # average duration of each launch of the transform: HIP events around `steps`
    # launches of that launch alone, on the stream the kernels run on
    pass_ms = [time_it(lambda i=i: f32.run_pass(i), steps, warmup) for i in range(f32.npasses)]
    # the roofline is the whole transform's: its algorithmic bytes (one read and
    # one write of the array) over the time of all its launches
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    traffic = [profiled_traffic(i) for i in range(f32.npasses)]
    out = {
        "metric": METRIC,
        "value": flops / (ms * 1e-3) / 1e9,
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: N(0,1) re/im from splitmix64 seed 1, resident in HBM",
        "config": {"workload": "batched C2C FFT, Complex Float, 64 x 65536, forward, planar layout",
                   "batch": batch, "length": n, "plan": f"{f32.npasses} passes"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None if None in traffic else int(sum(traffic)),
                     "algorithmic_bytes": alg_bytes,
                     "kernel": f"the whole 64 x 65536 transform: {f32.npasses} launch(es) of "
                               + ", ".join(launch_name(tune_report, i) for i in range(f32.npasses)),
                     "per_launch": [{"launch": launch_name(tune_report, i), "ms": pass_ms[i],
                                     "GBps_if_alone": alg_bytes / (pass_ms[i] * 1e-3) / 1e9,
                                     "traffic": traffic[i]} for i in range(f32.npasses)]},
    }
    if tune_report:
        out["config"]["autotune"] = [l for l in tune_report.strip().split("\n")]
    # a second denominator: the device's own device-to-device copy of the same
    # 32 MiB array (reads and writes 32 MiB each)
    src = torch.empty(batch * n * 2, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    msc = time_it(lambda: dst.copy_(src), steps, warmup)
    copy_gbps = alg_bytes / (msc * 1e-3) / 1e9
    out["roofline"]["device_copy_GBps"] = copy_gbps
    out["roofline"]["achieved_over_device_copy"] = achieved / copy_gbps
    del src, dst
    y32 = f32.result()

    if not args.skip_extras:
        extras = {}
        # FP32 error against the FP64 path and the oracle on sampled rows
        f64 = C2C(n, batch, _abi.F64)
        f64.set_io(x)
        ms64 = time_it(f64.run, steps, warmup)
        y64 = f64.result()
        extras["fp32_rel_l2_vs_fp64"] = rel_l2(y32, y64)
        extras["fp32_max_abs_vs_fp64"] = float(np.max(np.abs(y32 - y64)))
        extras["fp64_rel_l2_vs_numpy_2rows"] = rel_l2(y64[:2], np.fft.fft(x[:2], axis=-1))
        extras["fp64_same_workload_ms"] = ms64
        extras["fp32_over_fp64_speedup"] = ms64 / ms
        # external comparator (timing only): the vendor FFT behind torch.fft on ROCm,
        # same 64 x 65536 Complex Float, interleaved, out of place
        xt = torch.from_numpy(x).to("cuda", torch.complex64)
        yt = torch.empty_like(xt)
        msr = time_it(lambda: torch.fft.fft(xt, dim=-1, out=yt), steps, warmup)
        extras["comparator_torch_fft_rocfft_64x65536_f32_ms"] = msr
        extras["comparator_rel_l2_vs_this"] = rel_l2(yt.cpu().numpy(), y32)
        del xt, yt
        del f64
        # one 2^24 transform
        n24 = 1 << 24
        x24 = synthetic.complex_normal(1, (1, n24))
        g = C2C(n24, 1, _abi.F32)
        g.set_io(x24)
        if not args.no_tune:
            extras["c2b_autotune"] = g.tune().strip().split("\n")
        ms24 = time_it(g.run, max(steps // 2, 3), 2)
        b24 = 16.0 * n24
        extras["c2b_2p24_f32_ms"] = ms24
        extras["c2b_2p24_f32_GFLOPs"] = 5.0 * n24 * 24 / (ms24 * 1e-3) / 1e9
        extras["c2b_2p24_hbm_frac"] = b24 / (ms24 * 1e-3) / 1e9 / HBM_PEAK_GBS
        extras["c2b_roofline"] = {"bound": "hbm", "achieved": b24 / (ms24 * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": b24 / (ms24 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                  "traffic": profiled_c2b_traffic(), "algorithmic_bytes": int(b24),
                                  "kernel": f"the whole 2^24 transform: {g.npasses} launches"}
        y24 = g.result()
        extras["c2b_2p24_rel_l2_vs_numpy_fp64"] = rel_l2(y24[0], np.fft.fft(x24[0]))
        del g
        xt = torch.from_numpy(x24).to("cuda", torch.complex64)
        yt = torch.empty_like(xt)
        extras["comparator_torch_fft_rocfft_2p24_f32_ms"] = time_it(lambda: torch.fft.fft(xt, dim=-1, out=yt),
                                                                     max(steps // 2, 3), 2)
        del xt, yt
        # plain FP64, 256 x 1024
        x1 = synthetic.complex_uniform(0, (256, 1024))
        c1 = C2C(1024, 256, _abi.F64)
        c1.set_io(x1)
        if not args.no_tune:
            extras["c1_autotune"] = c1.tune().strip().split("\n")
        ms1 = time_it(c1.run, steps, warmup)
        extras["c1_f64_256x1024_us"] = ms1 * 1e3
        extras["c1_f64_GBps"] = 2 * 16 * 1024 * 256 / (ms1 * 1e-3) / 1e9
        # forward, then inverse, then scale by 1/N: the round trip and the
        # forward spectrum against numpy's FP64 FFT
        y1 = c1.result()
        c1i = C2C(1024, 256, _abi.F64, sign=1)
        c1i.set_io(y1)
        c1i.run()
        extras["c1_forward_rel_l2_vs_numpy"] = rel_l2(y1, np.fft.fft(x1, axis=-1))
        extras["c1_roundtrip_rel_l2"] = rel_l2(c1i.result() / 1024, x1)
        del c1, c1i
        xt = torch.from_numpy(x1).to("cuda", torch.complex128)
        yt = torch.empty_like(xt)
        extras["comparator_torch_fft_rocfft_c1_f64_us"] = time_it(lambda: torch.fft.fft(xt, dim=-1, out=yt),
                                                                   steps, warmup) * 1e3
        del xt, yt
        # real convolution and its adjoint, FP32 512 x 8192
        fb, N = conv_workload(torch)
        msc = time_it(fb, steps, warmup)
        extras["c4_conv_fwd_bwd_ms"] = msc
        extras["c4_conv_per_s"] = 512 / (msc * 1e-3)
        extras["c4_hbm_frac"] = 32 * 512 * N / (msc * 1e-3) / 1e9 / HBM_PEAK_GBS
        # external comparator (timing only): the same forward + gradient step with
        # torch.fft (rocFFT) and autograd on tensors of the same shape
        xt = torch.randn(512, N, device="cuda").requires_grad_()
        kt = (torch.randn(512, N, device="cuda") / math.sqrt(N)).requires_grad_()
        wt = torch.randn(512, N, device="cuda")

        def torch_fb():
            yt = torch.fft.irfft(torch.fft.rfft(xt) * torch.fft.rfft(kt), n=N) * N
            (yt * wt).sum().backward()
        extras["comparator_torch_fft_rocfft_c4_fwd_bwd_ms"] = time_it(torch_fb, steps, warmup)
        del xt, kt, wt
        # Karatsuba, 1024 pairs, lengths U[256, 4096]
        kar, lens, kdata = karatsuba_workload(torch, np)
        msk = time_it(kar, max(steps, 5), 2)
        extras["c5_karatsuba_ms"] = msk
        extras["c5_pairs_per_s"] = 1024 / (msk * 1e-3)
        extras["c5_schoolbook_equiv_GMACs"] = float(np.sum(lens.astype(np.float64) ** 2)) / (msk * 1e-3) / 1e9
        # compute roofline: the kernel's own leaf multiply-adds (3^d leaf products of
        # s x s coefficients per pair, P = 2^d s, exact in FP64 for these magnitudes)
        # against the FP64 vector peak (one FMA = 2 flops)
        macs = float(np.sum([karatsuba_subtree_macs(int(L)) for L in lens]))
        extras["c5_roofline"] = {"bound": "mfma-fp64", "achieved": 2 * macs / (msk * 1e-3) / 1e12,
                                 "peak": 78.6, "unit": "TFLOP/s",
                                 "frac": 2 * macs / (msk * 1e-3) / 1e12 / 78.6, "subtree_macs": macs,
                                 "kernel": "k_karatsuba_mfma (one workgroup per pair, FP64 MFMA subtree products; "
                                           "k_karatsuba skips the pairs it took)"}
        extras.update(karatsuba_fft_f32(torch, np, lens, kdata, steps, warmup))
        del kar, kdata
        extras.update(uneven_batch(torch, np, steps, tune=not args.no_tune))
        out["extras"] = extras

    if not args.skip_cpu:
        import oracle
        threads = oracle.num_threads()
        rows = 16
        xs = x[:rows]
        t0 = time.perf_counter()
        reps = 0
        while True:
            oracle.transform(-1, xs, precision="f32")
            reps += 1
            if time.perf_counter() - t0 > 10.0:
                break
        dt = (time.perf_counter() - t0) / reps
        out["cpu_baseline"] = {"value": 5.0 * n * math.log2(n) * rows / dt / 1e9, "unit": "GFLOP/s",
                               "cores": threads, "kind": "port",
                               "sample": f"{rows} of the 64 rows x 65536, FP32 oracle (C restatement of the "
                                         f"reference's split-radix plan), {reps} repetitions over ~10 s"}
        # a separate, labelled datum: scipy's pocketfft on the same rows and threads
        try:
            import scipy.fft as sfft
            x16 = xs.astype(np.complex64)
            sfft.fft(x16, axis=-1, workers=threads)
            t0 = time.perf_counter()
            reps = 0
            while time.perf_counter() - t0 < 2.0:
                sfft.fft(x16, axis=-1, workers=threads)
                reps += 1
            dt = (time.perf_counter() - t0) / reps
            out["cpu_baseline"]["scipy_pocketfft_f32_GFLOPs"] = 5.0 * n * math.log2(n) * rows / dt / 1e9
        except ImportError:
            pass
    print(json.dumps(out))


if __name__ == "__main__":
    main()
