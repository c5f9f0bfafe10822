#!/bin/sh
# Builds the CPU oracle (test infrastructure) into oracle/liborc.so.
# Plain IEEE arithmetic: no contraction, no fast-math, SSE2 scalar floats.
set -e
cd "$(dirname "$0")"
gcc -O2 -fPIC -shared -fopenmp -ffp-contract=off -fno-fast-math -std=c99 -Wall -Wno-unknown-pragmas \
    -o liborc.so orc.c
