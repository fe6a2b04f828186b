#!/bin/bash
# Builds libaf_dryrun.so: the library's sources compiled host-only and linked
# against fake_hip.cpp instead of the HIP runtime (kernels are checked, not run).
# Usage: tests/dryrun/build.sh [output-dir]   (default: tests/dryrun/_build)
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REPO="$(cd "$HERE/../.." && pwd)"
CSRC="$REPO/accelerate-fourier_amd/csrc"
OUT="${1:-$HERE/_build}"
mkdir -p "$OUT"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
FLAGS=(--cuda-host-only -O1 -g -std=c++17 -fPIC -I "$REPO/include" -I "$CSRC" -Wno-unused-command-line-argument)
objs=()
for f in fft_kernels.hip fast_pass.hip fast_ext.hip decomp.hip karatsuba.hip abi.cpp planner.cpp; do
    o="$OUT/${f%.*}.o"
    "$HIPCC" "${FLAGS[@]}" -x hip -c "$CSRC/$f" -o "$o" 2> "$OUT/${f%.*}.log" || { cat "$OUT/${f%.*}.log"; exit 1; }
    objs+=("$o")
done
# host-only objects still name the (absent) device code blobs: give them empty ones
: > "$OUT/fatbins.c"
for s in $(nm -u "${objs[@]}" | awk '/__hip_fatbin/ {print $2}' | sort -u); do
    echo "const char $s[16] = {0};" >> "$OUT/fatbins.c"
done
gcc -c -fPIC "$OUT/fatbins.c" -o "$OUT/fatbins.o"
g++ -O1 -g -std=c++17 -fPIC -I /opt/rocm/include -I "$CSRC" -D__HIP_PLATFORM_AMD__ -c "$HERE/fake_hip.cpp" -o "$OUT/fake_hip.o"
g++ -shared -Wl,-Bsymbolic -o "$OUT/libaf_dryrun.so" "${objs[@]}" "$OUT/fatbins.o" "$OUT/fake_hip.o"
echo "$OUT/libaf_dryrun.so"
