"""Launch checks without a GPU.

The library is built host-only against tests/dryrun/fake_hip.cpp: kernels do
not run, but every launch's memory footprint (from its pass descriptor or
scalar arguments) is checked against the live allocations, and its launch shape
against what the kernel assumes (thread count, LDS size, tile geometry).  Plans
are created and executed for lengths across every algorithm and for the
4096-row uneven batch, in both precisions.
"""
import ctypes
import importlib.util
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "dryrun", "_build")
LIB = os.path.join(BUILD, "libaf_dryrun.so")

pytestmark = pytest.mark.skipif(not os.path.exists(os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")),
                                reason="hipcc is needed to build the host-only library")

V = ctypes.c_void_p
C2C, R2C_FULL, C2R_FULL, CONV_CYCLIC, REAL_CONV = range(5)


class PlanDesc(ctypes.Structure):
    _fields_ = [("length", ctypes.c_int64), ("sign", ctypes.c_int32), ("kind", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("algorithm", ctypes.c_int32), ("flags", ctypes.c_uint32)]


@pytest.fixture(scope="module")
def dry():
    r = subprocess.run(["bash", os.path.join(HERE, "dryrun", "build.sh"), BUILD], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    lib = ctypes.CDLL(LIB)
    lib.dry_problem.restype = ctypes.c_char_p
    lib.dry_problem_count.restype = ctypes.c_long
    lib.dry_launches.restype = ctypes.c_long
    lib.dry_kernel_summary.restype = ctypes.c_char_p
    lib.hipMalloc.argtypes = [ctypes.POINTER(V), ctypes.c_size_t]
    lib.hipFree.argtypes = [V]
    lib.af_plan_create.argtypes = [ctypes.POINTER(PlanDesc), ctypes.POINTER(V)]
    lib.af_plan_destroy.argtypes = [V]
    lib.af_execute.argtypes = [V, ctypes.c_int64, V, V, V, V, V, V]
    lib.af_execute_noalloc.argtypes = [V, ctypes.c_int64, V, V, V, V, V]
    lib.af_plan_workspace_size.argtypes = [V, ctypes.c_int64, ctypes.POINTER(ctypes.c_size_t)]
    lib.af_vjp_execute.argtypes = [V, ctypes.c_int64, V, V, V, V, V, V]
    lib.af_to_spectrum.argtypes = [V, ctypes.c_int64, V, V, V, V, V]
    lib.af_from_spectrum.argtypes = [V, ctypes.c_int64, V, V, V, V, V]
    lib.af_convolve_cyclic.argtypes = [V, ctypes.c_int64, V, V, V, V, V, V, V, V]
    lib.af_real_convolve.argtypes = [V, ctypes.c_int64, V, V, V, V]
    lib.af_vjp_real_convolve.argtypes = [V, ctypes.c_int64, V, V, V, V, V, V]
    lib.af_batch_create.argtypes = [ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), ctypes.c_int32, ctypes.c_int32,
                                    ctypes.POINTER(V)]
    lib.af_batch_execute.argtypes = [V, V, V, V, V, V]
    lib.af_batch_destroy.argtypes = [V]
    lib.af_karatsuba_ragged_i64.argtypes = [ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), V, V, V, V]
    lib.af_karatsuba_i64.argtypes = [ctypes.c_int64, ctypes.c_int64, V, V, V, V]
    lib.af_karatsuba_cyclic_i64.argtypes = [ctypes.c_int64, ctypes.c_int64, V, V, V, V]
    lib.af_karatsuba_complex_i64.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int32] + [V] * 7
    return lib


class Bufs:
    def __init__(self, lib):
        self.lib, self.ptrs = lib, []

    def __call__(self, nbytes):
        p = V()
        assert self.lib.hipMalloc(ctypes.byref(p), max(int(nbytes), 1)) == 0
        self.ptrs.append(p)
        return p

    def free(self):
        for p in self.ptrs:
            self.lib.hipFree(p)
        self.ptrs = []


def problems(lib, start=0):
    n = lib.dry_problem_count()
    return [lib.dry_problem(i).decode() for i in range(start, n)]


def run_c2c(lib, n, sign, prec, batch, planar, alg=0):
    e = 16 if prec == 0 else 8
    d = PlanDesc(n, sign, C2C, prec, alg, 0)
    p = V()
    st = lib.af_plan_create(ctypes.byref(d), ctypes.byref(p))
    if st != 0:
        return st
    b = Bufs(lib)
    h = e // 2
    if planar:
        xr, xi, yr, yi = b(batch * n * h), b(batch * n * h), b(batch * n * h), b(batch * n * h)
    else:
        xr, xi, yr, yi = b(batch * n * e), None, b(batch * n * e), None
    st = lib.af_execute(p, batch, xr, xi, yr, yi, None, None)
    if st == 0:
        st = lib.af_vjp_execute(p, batch, xr, xi, yr, yi, None, None)
    b.free()
    lib.af_plan_destroy(p)
    return st


LENGTHS = sorted(set(list(range(0, 130)) + [240, 1155, 3003, 9240, 15015, 16016,  # prime-factor order
                                            255, 256, 257, 360, 1000, 1024, 2048, 4096, 4099, 5000, 8192, 16384,
                                            16385, 20000, 32768, 65521, 65536, 100003, 114689, 2 * 65521,
                                            1 << 20, 1200 * 1200, 997637]))


@pytest.mark.parametrize("prec", [0, 1])
def test_c2c_lengths(dry, prec):
    before = dry.dry_problem_count()
    for n in LENGTHS:
        for sign in (-1, 1):
            for batch, planar in ((1, False), (3, True), (64, False)):
                if n * batch > 1 << 22 and batch > 1:
                    continue
                st = run_c2c(dry, n, sign, prec, batch, planar)
                assert st in (0, 2), (n, sign, batch, planar, st)
    assert problems(dry, before) == []


def test_large_power_of_two(dry):
    before = dry.dry_problem_count()
    for n, batch in ((65536, 64), (1 << 24, 1), (1 << 24, 2), (1 << 25, 1), (1024, 256)):
        for prec in (0, 1):
            for planar in (False, True):
                assert run_c2c(dry, n, -1, prec, batch, planar) == 0
    assert problems(dry, before) == []


@pytest.mark.parametrize("prec", [0, 1])
def test_real_and_convolution(dry, prec):
    before = dry.dry_problem_count()
    e = 16 if prec == 0 else 8
    h = e // 2
    for n in (2, 4, 6, 10, 64, 86, 1000, 8192, 16384, 32768, 40000, 2 * 65521):
        for batch in (1, 5):
            b = Bufs(dry)
            for kind, fn in ((R2C_FULL, "to"), (C2R_FULL, "from")):
                d = PlanDesc(n, -1 if kind == R2C_FULL else 1, kind, prec, 0, 0)
                p = V()
                st = dry.af_plan_create(ctypes.byref(d), ctypes.byref(p))
                assert st in (0, 2), (n, kind, st)
                if st:
                    continue
                if fn == "to":
                    assert dry.af_to_spectrum(p, batch, b(batch * n * h), b(batch * n * e), None, None, None) == 0
                else:
                    assert dry.af_from_spectrum(p, batch, b(batch * n * e), None, b(batch * n * h), None, None) == 0
                dry.af_plan_destroy(p)
            d = PlanDesc(n, -1, REAL_CONV, prec, 0, 0)
            p = V()
            st = dry.af_plan_create(ctypes.byref(d), ctypes.byref(p))
            assert st in (0, 2), (n, st)
            if st == 0:
                x, k, y = b(batch * n * h), b(batch * n * h), b(batch * n * h)
                assert dry.af_real_convolve(p, batch, x, k, y, None) == 0
                assert dry.af_vjp_real_convolve(p, batch, x, k, b(batch * n * h), b(batch * n * h),
                                                b(batch * n * h), None) == 0
                dry.af_plan_destroy(p)
            b.free()
    for n in (1, 7, 97, 1000, 4099, 65521, 100003):
        b = Bufs(dry)
        d = PlanDesc(n, -1, CONV_CYCLIC, prec, 0, 0)
        p = V()
        assert dry.af_plan_create(ctypes.byref(d), ctypes.byref(p)) == 0
        assert dry.af_convolve_cyclic(p, 2, b(2 * n * e), None, b(2 * n * e), None, b(2 * n * e), None, None,
                                      None) == 0
        dry.af_plan_destroy(p)
        b.free()
    assert problems(dry, before) == []


def _uneven_lengths():
    spec = importlib.util.spec_from_file_location(
        "synthetic", os.path.join(REPO, "accelerate-fourier_amd", "accelerate_fourier", "synthetic.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m.uneven_lengths(2)[0]


@pytest.mark.parametrize("prec", [1, 0])
def test_uneven_batch_4096_rows(dry, prec):
    """The 4096-row uneven batch: every plan (with its plan-time FP64 transforms)
    and the captured grouped launches (one chunk of rows without intermediates,
    one with); in FP32 also,
    on every 7th row, every schedule the batch autotuner builds and launches:
    chunk sizes, and the rows without intermediates on a second captured stream
    (the stand-in runtime's events read 1 ms for every interval, so the first
    candidate, no chunking on one lane, is kept and launched again)."""
    before = dry.dry_problem_count()
    lens = _uneven_lengths()
    arr = (ctypes.c_int64 * len(lens))(*[int(x) for x in lens])
    h = V()
    assert dry.af_batch_create(len(lens), arr, -1, prec, ctypes.byref(h)) == 0
    e = 16 if prec == 0 else 8
    b = Bufs(dry)
    tot = int(np.sum(lens))
    dry.af_batch_describe.argtypes = [V, ctypes.c_char_p, ctypes.c_size_t]
    desc = ctypes.create_string_buffer(1 << 20)
    assert dry.af_batch_describe(h, desc, len(desc)) == 0
    head = desc.value.decode().split("\n")[0]
    assert "chunk_bytes=0" in head and int(head.split(" chunks=")[1].split()[0]) == 2, head
    assert dry.af_batch_execute(h, b(tot * e), None, b(tot * e), None, None) == 0
    dry.af_batch_destroy(h)
    if prec == 1:
        lens = lens[::7]
        arr = (ctypes.c_int64 * len(lens))(*[int(x) for x in lens])
        tot = int(np.sum(lens))
        assert dry.af_batch_create(len(lens), arr, -1, prec, ctypes.byref(h)) == 0
        dry.af_batch_tune.argtypes = [V, V, ctypes.c_char_p, ctypes.c_size_t]
        rep = ctypes.create_string_buffer(8192)
        assert dry.af_batch_tune(h, None, rep, len(rep)) == 0
        text = rep.value.decode()
        assert text.count("chunk_bytes=") == 9 and "chosen chunk_bytes=0 lanes=2" in text, text
        orders = [ln for ln in text.split("\n") if ln.startswith("launch ")]
        assert orders and all(" order=0 " in ln for ln in orders), text
        assert dry.af_batch_describe(h, desc, len(desc)) == 0
        assert " chunks=" in desc.value.decode().split("\n")[0]
        h2 = b(tot * e // 2)
        assert dry.af_batch_execute(h, b(tot * e // 2), h2, b(tot * e // 2), b(tot * e // 2), None) == 0
        dry.af_batch_destroy(h)
    b.free()
    assert problems(dry, before) == []


def test_karatsuba(dry):
    before = dry.dry_problem_count()
    b = Bufs(dry)
    # past 4608 (the LDS kernel's longest padded pair) the top levels split in global memory
    for L in (1, 2, 63, 64, 65, 127, 128, 129, 500, 512, 513, 1000, 2048, 2049, 4095, 4096, 4097, 4608, 4609,
              5000, 9217, 16384, 100000):
        for batch in (1, 3):
            x, y = b(batch * L * 8), b(batch * L * 8)
            assert dry.af_karatsuba_i64(batch, L, x, y, b(batch * (2 * L - 1) * 8), None) == 0
            assert dry.af_karatsuba_cyclic_i64(batch, L, x, y, b(batch * L * 8), None) == 0
            if L in (1, 65, 4096, 4608, 5000):
                xi, yi = b(batch * L * 8), b(batch * L * 8)
                for cyc, nz in ((0, 2 * L - 1), (1, L)):
                    assert dry.af_karatsuba_complex_i64(batch, L, cyc, x, xi, y, yi, b(batch * nz * 8),
                                                        b(batch * nz * 8), None) == 0
    lens = np.array([256, 4096, 1, 77, 3000, 4096, 2049, 5000, 20000, 4609], dtype=np.int64)
    offs = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    arr = (ctypes.c_int64 * len(offs))(*[int(v) for v in offs])
    tot = int(offs[-1])
    assert dry.af_karatsuba_ragged_i64(len(lens), arr, b(tot * 8), b(tot * 8), b((2 * tot - len(lens)) * 8), None) == 0
    b.free()
    assert problems(dry, before) == []


def test_autotune_candidates(dry):
    """af_plan_tune launches every candidate tile shape and radix order of each
    pass (all checked here); the stand-in runtime's events read 1 ms for every
    interval, so the planned shapes stay.  Wisdom imported by hand then makes
    af_execute launch other shapes, which are checked too, and shapes no kernel
    can run are ignored."""
    dry.af_plan_tune.argtypes = [V, ctypes.c_int64, ctypes.c_int32, V, ctypes.c_char_p, ctypes.c_size_t]
    dry.af_wisdom_export.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    dry.af_wisdom_import.argtypes = [ctypes.c_char_p]
    dry.af_plan_num_passes.argtypes = [V, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32)]
    before = dry.dry_problem_count()
    launches0 = dry.dry_launches()
    cases = [(65536, 1, 64, 1), (20000, 1, 3, 0), (1000, 0, 5, 1), (100003, 1, 2, 0), (1 << 24, 1, 1, 1),
             (4099, 0, 2, 0), (360, 1, 7, 1)]
    for n, prec, batch, planar in cases:
        d = PlanDesc(n, -1, C2C, prec, 0, 0)
        p = V()
        assert dry.af_plan_create(ctypes.byref(d), ctypes.byref(p)) == 0
        rep = ctypes.create_string_buffer(4096)
        assert dry.af_plan_tune(p, batch, planar, None, rep, len(rep)) == 0, (n, prec)
        cnt = ctypes.c_int32()
        assert dry.af_plan_num_passes(p, batch, ctypes.byref(cnt)) == 0
        assert rep.value.decode().count("pass ") == cnt.value
        dry.af_plan_destroy(p)
    assert dry.dry_launches() - launches0 > 200
    need = ctypes.c_size_t()
    assert dry.af_wisdom_export(None, 0, ctypes.byref(need)) == 0
    buf = ctypes.create_string_buffer(need.value)
    assert dry.af_wisdom_export(buf, len(buf), ctypes.byref(need)) == 0
    text = buf.value.decode()
    assert len(text.strip().split("\n")) == len(cases)
    assert "f32 0 65536 64 1 : 0/0 0/0" in text
    # other shapes by hand: 65536 planar batch 64 with 8- and 32-line tiles, and
    # 20000 with reversed radix orders and a 64-line tile that no kernel can run
    assert dry.af_wisdom_import(b"f32 0 65536 64 1 : 8/0 32/0\nf32 0 20000 3 0 : 2/1 64/2\n") == 0
    assert dry.af_wisdom_import(b"f32 0 65536 64 1 : 3/0\n") != 0
    assert dry.af_wisdom_import(b"garbage\n") != 0
    for n, batch, planar in ((65536, 64, True), (20000, 3, False)):
        assert run_c2c(dry, n, -1, 1, batch, planar) == 0
    assert problems(dry, before) == []


def test_two_for_one_packing(dry):
    """twoToSpectrum / twoFromSpectrum launches (the transform, the packing
    kernels) for several lengths, layouts and precisions."""
    dry.af_two_to_spectrum.argtypes = [V, ctypes.c_int64, V, V, V, V, V, V, V]
    dry.af_two_from_spectrum.argtypes = [V, ctypes.c_int64, V, V, V, V, V, V, V]
    dry.af_untangle_spectra.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, V, V, V, V, V, V, V]
    dry.af_entangle_spectra.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, V, V, V, V, V, V, V]
    before = dry.dry_problem_count()
    for n in (1, 2, 7, 97, 1000, 4099, 65536, 100003):
        for prec in (0, 1):
            e = 16 if prec == 0 else 8
            h = e // 2
            batch = 3
            d = PlanDesc(n, -1, C2C, prec, 0, 0)
            p = V()
            assert dry.af_plan_create(ctypes.byref(d), ctypes.byref(p)) == 0
            b = Bufs(dry)
            x, y = b(batch * n * h), b(batch * n * h)
            for planar in (False, True):
                if planar:
                    a = (b(batch * n * h), b(batch * n * h))
                    bb = (b(batch * n * h), b(batch * n * h))
                else:
                    a, bb = (b(batch * n * e), None), (b(batch * n * e), None)
                assert dry.af_two_to_spectrum(p, batch, x, y, a[0], a[1], bb[0], bb[1], None) == 0
                assert dry.af_two_from_spectrum(p, batch, a[0], a[1], bb[0], bb[1], x, y, None) == 0
                z = b(batch * n * e)
                assert dry.af_entangle_spectra(d.precision, batch, n, a[0], a[1], bb[0], bb[1], z, None, None) == 0
                assert dry.af_untangle_spectra(d.precision, batch, n, z, None, a[0], a[1], bb[0], bb[1], None) == 0
                # the untangling reads two positions per output: no aliasing
                assert dry.af_untangle_spectra(d.precision, batch, n, z, None, z, None, bb[0], bb[1], None) != 0
            b.free()
            dry.af_plan_destroy(p)
    assert problems(dry, before) == []


def test_multidimensional(dry):
    """2-D and 3-D transforms (strided passes and transposes) and their two-for-one
    packing, every launch checked."""
    dry.af_execute_2d.argtypes = [V, V, ctypes.c_int64, V, V, V, V, V]
    dry.af_execute_3d.argtypes = [V, V, V, ctypes.c_int64, V, V, V, V, V]
    dry.af_two_to_spectrum_2d.argtypes = [V, V, ctypes.c_int64, V, V, V, V, V, V, V]
    dry.af_two_from_spectrum_2d.argtypes = [V, V, ctypes.c_int64, V, V, V, V, V, V, V]
    before = dry.dry_problem_count()

    def plan(n, prec):
        p = V()
        assert dry.af_plan_create(ctypes.byref(PlanDesc(n, -1, C2C, prec, 0, 0)), ctypes.byref(p)) == 0
        return p

    for prec in (0, 1):
        e = 16 if prec == 0 else 8
        for h, w in ((1, 1), (7, 5), (256, 1024), (4099, 16), (16, 4099), (65536, 2), (3, 65536), (1000, 3)):
            ph, pw = plan(h, prec), plan(w, prec)
            b = Bufs(dry)
            for planar in (False, True):
                if planar:
                    io = (b(2 * h * w * e // 2), b(2 * h * w * e // 2), b(2 * h * w * e // 2), b(2 * h * w * e // 2))
                else:
                    io = (b(2 * h * w * e), None, b(2 * h * w * e), None)
                assert dry.af_execute_2d(pw, ph, 2, *io, None) == 0
                assert dry.af_execute_2d(pw, ph, 2, io[2], io[3], io[2], io[3], None) == 0  # in place
            x, y = b(2 * h * w * e // 2), b(2 * h * w * e // 2)
            a, bb = b(2 * h * w * e), b(2 * h * w * e)
            assert dry.af_two_to_spectrum_2d(pw, ph, 2, x, y, a, None, bb, None, None) == 0
            assert dry.af_two_from_spectrum_2d(pw, ph, 2, a, None, bb, None, x, y, None) == 0
            b.free()
            dry.af_plan_destroy(ph)
            dry.af_plan_destroy(pw)
        for k, m, n in ((2, 3, 5), (3, 4099, 2), (64, 32, 48)):
            p0, p1, p2 = plan(n, prec), plan(m, prec), plan(k, prec)
            b = Bufs(dry)
            xin, xout = b(2 * k * m * n * e), b(2 * k * m * n * e)
            assert dry.af_execute_3d(p0, p1, p2, 2, xin, None, xout, None, None) == 0
            b.free()
            for p in (p0, p1, p2):
                dry.af_plan_destroy(p)
    assert problems(dry, before) == []


def test_execute_noalloc(dry):
    """af_execute_noalloc: plans with scratch (four-step, Bluestein, Rader,
    three-pass) run their launches on the fixed per-device buffer (every
    footprint inside it); a call whose scratch exceeds the buffer returns
    AF_UNSUPPORTED and launches nothing."""
    before = dry.dry_problem_count()
    b = Bufs(dry)
    for n, prec, batch in ((65536, 1, 64), (65536, 0, 8), (100003, 1, 4), (4099, 0, 3), (65521, 1, 2),
                           (1 << 24, 1, 1), (1000, 0, 5)):
        d = PlanDesc(n, -1, C2C, prec, 0, 0)
        p = V()
        assert dry.af_plan_create(ctypes.byref(d), ctypes.byref(p)) == 0
        ws = ctypes.c_size_t()
        assert dry.af_plan_workspace_size(p, batch, ctypes.byref(ws)) == 0
        e = 16 if prec == 0 else 8
        x, y = b(batch * n * e), b(batch * n * e)
        launches = dry.dry_launches()
        st = dry.af_execute_noalloc(p, batch, x, None, y, None, None)
        if ws.value <= 64 << 20:
            assert st == 0, (n, prec, batch)
            assert dry.dry_launches() > launches
        else:
            assert st == 2 and dry.dry_launches() == launches, (n, prec, batch)
        dry.af_plan_destroy(p)
    b.free()
    assert problems(dry, before) == []


def test_decompose_plans(dry):
    """transformDecompose of lengths with prime factors no single device plan
    serves (nested Rader, Good-Thomas and Cooley-Tukey compositions): every
    launch of the composition, sub-plans included, in both precisions, both
    layouts and both signs, with batches of rows."""
    before = dry.dry_problem_count()
    for n in (34, 47, 94, 289, 4418, 10007):
        for prec in (0, 1):
            for sign, batch, planar in ((-1, 1, False), (1, 3, True)):
                assert run_c2c(dry, n, sign, prec, batch, planar, alg=1) == 0, (n, prec, sign)
    assert problems(dry, before) == []

