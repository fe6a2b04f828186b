#!/bin/sh
# Builds the CPU oracle (test infrastructure only) into oracle/liborc.so.
# The reference itself is Haskell on the Accelerate EDSL and cannot be built
# here (no GHC toolchain), so there is no oracle/_ref recipe.
set -e
cd "$(dirname "$0")"
gcc -O2 -std=c11 -ffp-contract=off -fno-fast-math -fopenmp -fPIC -shared \
    -Wall -Wextra -Wno-unused-parameter -Wno-maybe-uninitialized \
    -o liborc.so fourier_oracle.c -lm
