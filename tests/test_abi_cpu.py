"""The C-ABI boundary without a GPU: the library loads, exports every symbol
the header declares, validates every precondition the reference documents but
leaves unchecked, plans on the host, and its host-side index maps agree bit for
bit with the oracle's restatement of the reference (Permutation.hs,
Planned.hs:597-617, NumberTheory.hs).  No compute call is made here.
"""
import ctypes
import re

import numpy as np
import pytest

from accelerate_fourier import _abi, planned
from accelerate_fourier.sign import Sign

L = None


@pytest.fixture(scope="module")
def lib():
    return _abi.lib()


def header_symbols():
    text = open(_abi.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(af_[a-z0-9_]+)\s*\(", text)))


def test_every_header_symbol_is_exported(lib):
    syms = header_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), f"{s} declared in the header but not exported"
    assert sorted(_abi.exported_symbols()) == syms, "ctypes table and header disagree"


def test_status_strings(lib):
    for st in range(7):
        assert len(_abi.status_string(st)) > 3
    assert "unknown" in _abi.status_string(7)
    assert "unknown" in _abi.status_string(99)
    assert lib.af_version().decode().startswith("accelerate-fourier-amd")


def desc(length=8, sign=-1, kind=_abi.C2C, prec=_abi.F32, alg=_abi.ALG_AUTO, flags=0):
    return _abi.PlanDesc(length, sign, kind, prec, alg, flags)


def create(lib, d):
    h = ctypes.c_void_p()
    return lib.af_plan_create(ctypes.byref(d), ctypes.byref(h))


@pytest.mark.parametrize("d", [
    desc(length=-1),
    desc(sign=0), desc(sign=2), desc(sign=-2),
    desc(kind=9), desc(prec=7), desc(alg=5), desc(flags=0x80),
    desc(length=7, kind=_abi.R2C_FULL),      # Real.hs:45-47: even size
    desc(length=9, kind=_abi.C2R_FULL),      # Real.hs:88-90
    desc(length=0, kind=_abi.CONV_CYCLIC),   # Planned.hs:776-778: non-empty
    desc(length=0, kind=_abi.REAL_CONV),
    desc(length=5, kind=_abi.REAL_CONV),
])
def test_invalid_descriptors(lib, d):
    assert create(lib, d) == _abi.INVALID_ARG
    buf = ctypes.create_string_buffer(256)
    assert lib.af_plan_describe(ctypes.byref(d), buf, len(buf)) == _abi.INVALID_ARG


def test_null_arguments(lib):
    d = desc()
    assert lib.af_plan_create(ctypes.byref(d), None) == _abi.INVALID_ARG
    assert lib.af_plan_create(None, ctypes.byref(ctypes.c_void_p())) == _abi.INVALID_ARG
    assert lib.af_plan_destroy(None) == _abi.INVALID_ARG
    assert lib.af_execute(None, 1, 1, None, 1, None, None, None) == _abi.INVALID_ARG
    assert lib.af_to_spectrum(None, 1, 1, 1, None, None, None) == _abi.INVALID_ARG
    assert lib.af_from_spectrum(None, 1, 1, None, 1, None, None) == _abi.INVALID_ARG
    assert lib.af_convolve_cyclic(None, 1, 1, None, 1, None, 1, None, None, None) == _abi.INVALID_ARG
    assert lib.af_real_convolve(None, 1, 1, 1, 1, None) == _abi.INVALID_ARG
    assert lib.af_vjp_execute(None, 1, 1, None, 1, None, None, None) == _abi.INVALID_ARG
    assert lib.af_vjp_real_convolve(None, 1, 1, 1, 1, 1, 1, None) == _abi.INVALID_ARG
    assert lib.af_batch_destroy(None) == _abi.INVALID_ARG
    assert lib.af_batch_execute(None, 1, None, 1, None, None) == _abi.INVALID_ARG
    n = ctypes.c_int32()
    assert lib.af_plan_num_passes(None, 1, ctypes.byref(n)) == _abi.INVALID_ARG


def test_karatsuba_preconditions(lib):
    # negative sizes, null buffers, unequal (empty) pair lengths -- checked before any device call
    assert lib.af_karatsuba_i64(-1, 4, 1, 1, 1, None) == _abi.INVALID_ARG
    assert lib.af_karatsuba_i64(1, -4, 1, 1, 1, None) == _abi.INVALID_ARG
    assert lib.af_karatsuba_i64(1, 4, None, 1, 1, None) == _abi.INVALID_ARG
    assert lib.af_karatsuba_i64(0, 4, None, None, None, None) == _abi.SUCCESS
    assert lib.af_karatsuba_cyclic_i64(1, 4, 1, None, 1, None) == _abi.INVALID_ARG
    off = _abi.int64_array([0, 3, 3, 7])  # second pair empty
    p = ctypes.cast(off, ctypes.POINTER(ctypes.c_int64))
    assert lib.af_karatsuba_ragged_i64(3, p, 1, 1, 1, None) == _abi.INVALID_ARG
    off2 = _abi.int64_array([1, 3])  # does not start at 0
    assert lib.af_karatsuba_ragged_i64(1, ctypes.cast(off2, ctypes.POINTER(ctypes.c_int64)), 1, 1, 1,
                                       None) == _abi.INVALID_ARG


def test_batch_preconditions(lib):
    lens = _abi.int64_array([4, -1])
    h = ctypes.c_void_p()
    assert lib.af_batch_create(2, ctypes.cast(lens, ctypes.POINTER(ctypes.c_int64)), -1, _abi.F32,
                               ctypes.byref(h)) == _abi.INVALID_ARG
    lens = _abi.int64_array([4])
    assert lib.af_batch_create(1, ctypes.cast(lens, ctypes.POINTER(ctypes.c_int64)), 3, _abi.F32,
                               ctypes.byref(h)) == _abi.INVALID_ARG


def test_valid_plan_without_device(lib):
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    assert create(lib, desc(1024)) == _abi.NO_DEVICE


def test_host_planner_descriptions():
    assert planned.describe(1024, precision=_abi.F64) == "c2c.f64:stockham1024[16,16,4]"
    assert planned.describe(65536) == "c2c.f32:fourstep65536(256[16,16]x256[16,16])"
    assert planned.describe(1 << 23).startswith("c2c.f32:fourstep8388608(2048")
    assert planned.describe(1 << 24) == "c2c.f32:threepass16777216(256[16,16]x256[16,16]x256[16,16])"
    # three passes for smooth lengths that two LDS passes cannot hold
    assert planned.describe(3 ** 18).startswith("c2c.f32:threepass387420489(729")
    # ... but never for the convolution cores
    assert planned.describe(1 << 24, kind=_abi.CONV_CYCLIC).startswith("convcyclic.f32:fourstep")
    assert planned.describe(97).startswith("c2c.f32:rader97(stockham96")
    assert planned.describe(65521).startswith("c2c.f32:rader65521(fourstep65520")
    assert planned.describe(100003).startswith("c2c.f32:bluestein100003(")
    assert planned.describe(47, algorithm=_abi.ALG_CHIRP2) == "c2c.f32:bluestein47(stockham128[16,8])"
    assert planned.describe(47, algorithm=_abi.ALG_CHIRP235) == "c2c.f32:bluestein47(stockham96[16,2,3])"
    assert planned.describe(0) == "c2c.f32:empty"
    assert planned.describe(1) == "c2c.f32:copy1"
    assert planned.describe(8192, kind=_abi.REAL_CONV) == "realconv.f32:stockham4096[16,16,16]"
    # single passes over pairwise coprime radices run in prime-factor order
    # (Good-Thomas, Planned.hs:584-617) unless Cooley-Tukey is asked for;
    # repeated primes (45 = 5 x 9, 4000 = 32 x 125) and convolutions keep it
    assert planned.describe(30) == "c2c.f32:primefactor30[2,5,3]"
    assert planned.describe(15) == "c2c.f32:stockham15[5,3]"  # two factors: Cooley-Tukey is faster
    assert planned.describe(15015) == "c2c.f32:primefactor15015[13,11,7,5,3]"
    assert planned.describe(1155, precision=_abi.F64) == "c2c.f64:primefactor1155[11,7,5,3]"
    assert planned.describe(60) == "c2c.f32:primefactor60[4,5,3]"
    assert planned.describe(30, flags=_abi.FLAG_COOLEY_TUKEY) == "c2c.f32:stockham30[2,5,3]"
    assert planned.describe(45) == "c2c.f32:stockham45[5,3,3]"
    assert planned.describe(4000) == "c2c.f32:stockham4000[16,2,5,5,5]"
    assert planned.describe(60, kind=_abi.CONV_CYCLIC) == "convcyclic.f32:stockham60[4,5,3]"
    assert planned.describe(61).startswith("c2c.f32:rader61(stockham60[4,5,3])")
    # a prime whose p-1 has a large factor: Rader would nest, so Bluestein; and
    # transformDecompose nests Rader through host-composed decompositions
    # (10006 = 2 x 5003: Good-Thomas; 5003 prime: Rader over 5002 = 82 x 61 ...)
    assert planned.describe(10007).startswith("c2c.f32:bluestein10007(")
    assert planned.describe(10007, algorithm=_abi.ALG_DECOMPOSE) == (
        "c2c.f32:rader10007(goodthomas10006(stockham2[2],rader5003(goodthomas5002(goodthomas82(stockham2[2],"
        "rader41(stockham40[8,5])),rader61(stockham60[4,5,3])))))")
    assert planned.describe(47, algorithm=_abi.ALG_DECOMPOSE) == (
        "c2c.f32:rader47(goodthomas46(stockham2[2],rader23(stockham22[2,11])))")
    assert planned.describe(289, algorithm=_abi.ALG_DECOMPOSE) == (
        "c2c.f32:cooleytukey289(rader17(stockham16[16]),rader17(stockham16[16]))")
    assert "bluestein" not in planned.describe(1000003, algorithm=_abi.ALG_DECOMPOSE)


def test_bluestein_pads_follow_the_reference(orc):
    """transformChirp235 pads to ceiling5Smooth(2n-1) and transformChirp2 to
    ceilingPowerOfTwo(2n-1) (Planned.hs:704-737), as the oracle's planner."""
    for n in (3, 47, 1000, 4099, 100003):
        s235 = planned.describe(n, algorithm=_abi.ALG_CHIRP235)
        s2 = planned.describe(n, algorithm=_abi.ALG_CHIRP2)
        m235 = int(re.search(r"\((?:stockham|fourstep)(\d+)", s235).group(1))
        m2 = int(re.search(r"\((?:stockham|fourstep)(\d+)", s2).group(1))
        assert m235 == orc.ceiling_5smooth(2 * n - 1)
        assert m2 == orc.ceiling_power_of_two(2 * n - 1)
        assert f"chirp{n}(" in orc.plan_string(n, orc.CHIRP235) and str(m235) in orc.plan_string(n, orc.CHIRP235)


def test_rader_permutations_bit_exact(lib, orc):
    for p in (3, 5, 7, 13, 17, 97, 257, 65521):
        perm = np.zeros(p - 1, np.int64)
        rev = np.zeros(p - 1, np.int64)
        inv = np.zeros(p - 1, np.int64)
        P = ctypes.POINTER(ctypes.c_int64)
        st = lib.af_rader_permutations(p, perm.ctypes.data_as(P), rev.ctypes.data_as(P), inv.ctypes.data_as(P))
        assert st == _abi.SUCCESS
        operm = orc.perm_multiplicative(p)
        assert np.array_equal(perm, operm)
        assert np.array_equal(rev, orc.perm_reverse(operm))
        assert np.array_equal(inv, orc.perm_inverse(operm))
    assert lib.af_rader_permutations(15, None, None, None) == _abi.INVALID_ARG


def test_good_thomas_maps_bit_exact(lib, orc):
    P = ctypes.POINTER(ctypes.c_int64)
    for n, m in ((3, 5), (5, 9), (7, 9), (63, 65), (4, 15)):
        a = np.zeros(n * m, np.int64)
        b = np.zeros(n * m, np.int64)
        assert lib.af_good_thomas_maps(n, m, a.ctypes.data_as(P), b.ctypes.data_as(P)) == _abi.SUCCESS
        assert np.array_equal(a, orc.skew_grid(m, n))
        assert np.array_equal(b, orc.skew_grid_inv(n, m))
    a = np.zeros(36, np.int64)
    assert lib.af_good_thomas_maps(6, 6, a.ctypes.data_as(P), a.ctypes.data_as(P)) == _abi.INVALID_ARG


def test_number_theory_matches_oracle(lib, orc):
    for p in (3, 5, 7, 11, 97, 257, 65521, 100003, 1000003):
        assert lib.af_multiplicative_generator(p) == orc.multiplicative_generator(p)
    assert lib.af_multiplicative_generator(15) == -1
    for n in list(range(1, 300)) + [199999, 2000005]:
        assert lib.af_ceiling_5smooth(n) == orc.ceiling_5smooth(n)
        assert lib.af_ceiling_power_of_two(n) == orc.ceiling_power_of_two(n)


def test_planners_and_caches_host(orc):
    """Planned.plan / planWithMapUpdate / planDecomposeWithMapUpdate /
    planChirpWithMapUpdate (Planned.hs:201-293, 731-737) select the device
    planner's algorithms; the chirp pads follow the oracle's planner, and
    transformWithPlanner / cacheDuplex build transforms of the right sign and
    algorithm without a device."""
    assert planned.plan(1024) == planned.describe(1024)
    assert planned.plan_with_map_update(10007).startswith("c2c.f32:bluestein10007(")
    assert planned.plan_decompose_with_map_update(10007).startswith("c2c.f32:rader10007(goodthomas10006(")
    assert planned.plan_decompose_with_map_update(65521).startswith("c2c.f32:rader65521(")
    for pad, alg in ((planned.ceiling_power_of_two, orc.CHIRP2), (planned.ceiling_5_smooth, orc.CHIRP235)):
        for n in (3, 47, 1000):
            s = planned.plan_chirp_with_map_update(pad)(n)
            m = int(re.search(r"\((?:stockham|fourstep)(\d+)", s).group(1))
            assert m == pad(2 * n - 1)
            assert str(m) in orc.plan_string(n, alg)
    with pytest.raises(ValueError):
        planned.plan_chirp_with_map_update(lambda n: n)
    with pytest.raises(TypeError):
        planned.transform_with_planner(lambda n: None, Sign(-1), 8)
    t = planned.transform_with_planner(planned.plan_decompose_with_map_update, Sign(1), 97)
    assert (t.sign.value, t.n, t.algorithm) == (1, 97, _abi.ALG_DECOMPOSE)
    ci, cf = planned.cache_duplex(12)
    ti, tf = planned.transform_with_cache(ci), planned.transform_with_cache(cf)
    assert (ti.sign.value, tf.sign.value, ti.n, tf.n) == (1, -1, 12, 12)
    assert planned.transform_with_cache(planned.cache(Sign(1), 5)).sign.value == 1
