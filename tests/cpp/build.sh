#!/bin/sh
# Builds the C++ shim clients (tests/cpp/_build/shim_filter, shim_ode) with the host
# compiler only: the shims reach the GPU through include/bihip.h and
# libbi_amd/lib/libbihip.so.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
mkdir -p "$HERE/_build"
${CXX:-g++} -std=c++17 -O2 -Wall -Wextra -I "$ROOT/include" "$HERE/shim_filter.cpp" \
  -L "$ROOT/libbi_amd/lib" -lbihip -Wl,-rpath,'$ORIGIN/../../../libbi_amd/lib' \
  -o "$HERE/_build/shim_filter.tmp"
mv "$HERE/_build/shim_filter.tmp" "$HERE/_build/shim_filter"
${CXX:-g++} -std=c++17 -O2 -Wall -Wextra -I "$ROOT/include" "$HERE/shim_ode.cpp" \
  -L "$ROOT/libbi_amd/lib" -lbihip -Wl,-rpath,'$ORIGIN/../../../libbi_amd/lib' \
  -o "$HERE/_build/shim_ode.tmp"
mv "$HERE/_build/shim_ode.tmp" "$HERE/_build/shim_ode"
