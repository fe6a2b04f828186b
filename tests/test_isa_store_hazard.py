"""The generated gfx950 code of every kernel source keeps a wait state between
a vector-memory store of more than 8 bytes and a vector instruction that
overwrites the store's data registers, and spills nothing to scratch.

A buffer store that takes its wavefront-uniform offset from an SGPR (soffset)
gets no such wait state from the compiler, and on MI355X those stores then
wrote partly overwritten data under load (FP64 interleaved 64 x 65536
transforms came out wrong from about 32 rows on).  The kernels keep 16-byte
stores on VGPR-only offsets; this test compiles the sources to assembly (no
GPU needed) and fails if any such store is directly followed by a write of its
data registers.
"""
import concurrent.futures as cf
import os
import re
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(REPO, "accelerate-fourier_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["fast_pass.hip", "fast_ext.hip", "fft_kernels.hip", "karatsuba.hip"]

pytestmark = pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc is needed to compile the kernels")

_STORE = re.compile(r"^(buffer|global|flat)_store_dwordx(3|4)\s+(.*)$")
_VREG = re.compile(r"v\[(\d+):(\d+)\]|v(\d+)\b")
_VALU_DST = re.compile(r"^v_\S+\s+(v\[(\d+):(\d+)\]|v(\d+)\b)")


def _range(m):
    if m.group(1) is not None:
        return int(m.group(1)), int(m.group(2))
    return int(m.group(3)), int(m.group(3))


def _asm(src, out_dir):
    out = os.path.join(out_dir, os.path.splitext(src)[0] + ".s")
    cmd = [HIPCC, "-O3", "-std=c++17", "--offload-arch=gfx950", "--cuda-device-only", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "-I", os.path.join(REPO, "include"), "-I", CSRC, "-S", "-c",
           os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return out


def clobbered_stores(asm_text):
    """(kernel, store, next instruction) for every store of more than 8 bytes
    whose data registers the directly following vector instruction writes."""
    found = []
    for m in re.finditer(r"^(_Z\w+):", asm_text, re.M):
        end = asm_text.find(".Lfunc_end", m.end())
        body = [l.strip() for l in asm_text[m.end():end].split("\n")]
        body = [l for l in body if l and not l.startswith((";", ".")) and not l.endswith(":")]
        for k, line in enumerate(body[:-1]):
            st = _STORE.match(line)
            if not st:
                continue
            ops = [o.strip() for o in st.group(3).split(",")]
            data = ops[0] if st.group(1) == "buffer" else ops[1]
            dm = _VREG.match(data)
            if not dm:
                continue
            lo, hi = _range(dm)
            w = _VALU_DST.match(body[k + 1])
            if not w:
                continue
            a, b = _range(_VREG.match(w.group(1)))
            if not (b < lo or a > hi):
                found.append((m.group(1), line, body[k + 1]))
    return found


def test_scanner_finds_a_planted_clobber():
    """The scan itself: a store directly followed by a write of its data."""
    text = ("_Zk:\n\tbuffer_store_dwordx4 v[64:67], v124, s[8:11], s20 offen\n"
            "\tv_mul_f64 v[64:65], v[128:129], v[70:71]\n"
            "\tbuffer_store_dwordx4 v[0:3], v9, s[8:11], 0 offen\n\ts_nop 1\n\tv_mov_b32 v1, 0\n.Lfunc_end0:\n")
    hits = clobbered_stores(text)
    assert len(hits) == 1 and "v[64:67]" in hits[0][1]


_ASM_TEXTS = {}


def _asm_texts():
    """The gfx950 assembly of every kernel source, compiled once per session."""
    if not _ASM_TEXTS:
        with tempfile.TemporaryDirectory() as d:
            with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
                outs = list(ex.map(lambda s: _asm(s, d), SOURCES))
            for src, o in zip(SOURCES, outs):
                _ASM_TEXTS[src] = open(o).read()
    return _ASM_TEXTS


def test_no_store_data_clobber():
    hits = []
    for text in _asm_texts().values():
        hits += clobbered_stores(text)
    assert not hits, f"{len(hits)} stores with clobbered data, e.g. {hits[:3]}"


# The FP64 generic pass kernels keep a few registers' worth of spills at the
# 128-register budget of 1024-thread workgroups; no other kernel may use
# scratch (private memory traffic per lane, paid in every workgroup).
_SCRATCH_ALLOWED = {"_ZN2af6k_passIdLb0EEEv6AfPass": 64, "_ZN2af6k_passIdLb1EEEv6AfPass": 64,
                    "_ZN2af14k_pass_groupedIdLb0EEEvPK6AfPassPKli": 64,
                    "_ZN2af14k_pass_groupedIdLb1EEEvPK6AfPassPKli": 64}


def test_no_kernel_spills():
    over = []
    for text in _asm_texts().values():
        for m in re.finditer(r"\.amdhsa_kernel (\S+)\n(.*?)\.end_amdhsa_kernel", text, re.S):
            size = re.search(r"\.amdhsa_private_segment_fixed_size (\d+)", m.group(2))
            if size and int(size.group(1)) > _SCRATCH_ALLOWED.get(m.group(1), 0):
                over.append((m.group(1), int(size.group(1))))
    assert not over, over
