#!/bin/sh
# Builds the CPU oracle (test infrastructure) into oracle/_build/.
# Flags: the reference's Release flags (CMakeLists.txt:49) plus -ffp-contract=off.
set -e
here="$(cd "$(dirname "$0")" && pwd)"
mkdir -p "$here/_build"
g++ -O3 -msse4.2 -ffp-contract=off -fno-fast-math -fopenmp -std=gnu++17 -fPIC -shared \
    -Wall -Wno-unused-function \
    -o "$here/_build/libnicp_oracle.so.tmp" "$here/nicp_oracle.cpp"
mv "$here/_build/libnicp_oracle.so.tmp" "$here/_build/libnicp_oracle.so"
