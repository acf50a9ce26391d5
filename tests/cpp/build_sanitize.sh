#!/bin/bash
# Host-only build of the planner and the CPU oracle with AddressSanitizer and
# UndefinedBehaviorSanitizer (g++; no GPU code is involved).
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(dirname "$(dirname "$HERE")")"
mkdir -p "$HERE/_build"
g++ -O1 -g -std=c++17 -fopenmp -pthread -fsanitize=address,undefined -fno-omit-frame-pointer \
    -I"$ROOT/fmca_amd/csrc" "$HERE/sanitize_check.cpp" "$ROOT/fmca_amd/csrc/plan.cpp" \
    "$ROOT/oracle/h2_oracle.cpp" -o "$HERE/_build/sanitize_check"
