#!/bin/sh
# Compiles the one reference source file on this path that builds from itself alone,
# MurmurHash2_64.cu (host-only C++; no CUDA, no moderngpu), straight from where it lies
# in the reference checkout, into oracle/_ref/libmurmur_ref.so.  The checker uses it to
# validate orc_murmur64a.  The rest of the reference path (cm.cu, filter.cu, select.cu,
# operators.cu ...) needs nvcc, the CUDA runtime and the absent moderngpu submodule, so
# it is unbuildable here.  Nothing is built when the reference is absent (the GPU box).
set -e
REF=${ALENKA_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
if [ ! -f "$REF/MurmurHash2_64.cu" ]; then
  echo "build_ref: reference not present at $REF; skipping" >&2
  exit 0
fi
mkdir -p "$HERE/_ref"
# nvcc pre-includes its runtime headers, which declare size_t; -include cstddef (a
# standard header of this image) stands in for that implicit include.
g++ -O2 -fPIC -shared -include cstddef -x c++ "$REF/MurmurHash2_64.cu" -o "$HERE/_ref/libmurmur_ref.so"
echo "build_ref: built $HERE/_ref/libmurmur_ref.so"
