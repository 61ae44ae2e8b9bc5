#!/bin/sh
# Regenerates tests/golden/philox4x32_10_rocrand.json from rocRAND's engine (CPU only).
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=${TMPDIR:-/tmp}/make_philox_vectors
/opt/rocm/bin/hipcc -O2 -x c++ -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include "$HERE/make_philox_vectors.cpp" -o "$OUT" 2>/dev/null || \
  /opt/rocm/bin/hipcc -O2 -I/opt/rocm/include "$HERE/make_philox_vectors.cpp" -o "$OUT"
"$OUT" > "$HERE/philox4x32_10_rocrand.json"
echo "wrote $HERE/philox4x32_10_rocrand.json"
