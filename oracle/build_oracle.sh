#!/bin/sh
# Builds the CPU oracle (test infrastructure only) into oracle/libpetrbf_oracle.so.
# -ffp-contract=off keeps every a*b+c as two roundings, as the reference's x86
# build without FMA evaluates it.
set -e
here="$(cd "$(dirname "$0")" && pwd)"
g++ -std=c++17 -O2 -fopenmp -ffp-contract=off -fPIC -shared \
    -o "$here/libpetrbf_oracle.so.tmp" "$here/petrbf_oracle.cpp"
mv "$here/libpetrbf_oracle.so.tmp" "$here/libpetrbf_oracle.so"
