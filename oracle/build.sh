#!/bin/sh
# Builds the oracle (test infrastructure): liboracle.so and evp_oracle.
# gcc, x86-64 baseline ISA (no FMA contraction), OpenMP for voxel loops.
set -e
cd "$(dirname "$0")"
CFLAGS="-O2 -fopenmp -fPIC -std=gnu99 -Wall -Wno-unused-result -Wno-stringop-overflow -Wno-stringop-overread"
gcc $CFLAGS -shared -o liboracle.so oracle.c -lm
gcc $CFLAGS -o evp_oracle evp_oracle_main.c oracle.c -lm
