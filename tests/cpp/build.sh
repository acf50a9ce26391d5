#!/bin/bash
# Builds the C++ facade check against the in-tree C-ABI library.
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(dirname "$(dirname "$HERE")")"
mkdir -p "$HERE/_build"
g++ -O2 -std=c++17 -Wall -I"$ROOT/include" "$HERE/facade_check.cpp" -o "$HERE/_build/facade_check" \
    -L"$ROOT/fmca_amd" -lfmca_amd -Wl,-rpath,"$ROOT/fmca_amd"
