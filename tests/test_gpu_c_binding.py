"""The reference-side binding's call sequence, replayed in C against the
library (tests/c_binding/replay.c, built by accelerate-fourier_amd/build.py):
plan creation per partial application, planar component arrays staged through
grow-only device buffers, af_execute / af_to_spectrum / af_from_spectrum /
af_karatsuba_i64 on the default stream, the device-resident forward -> inverse
chain, and the refused odd real length, for Double and Float.  Checked inside
the program against long-double DFTs and exact Int64 schoolbook products."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "c_binding", "replay")


def test_c_binding_replay():
    assert os.path.exists(BIN), "build first: python accelerate-fourier_amd/build.py"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().split("\n")
    assert lines[-1] == "replay ok"
    checks = [l for l in lines if " err " in l and l.endswith(" ok")]
    # per precision: 6 transforms, the device-resident chain, toSpectrum and
    # fromSpectrum; then 3 Karatsuba and the refused odd length
    assert len(checks) == 2 * 9 + 3 + 1, r.stdout
