#!/bin/sh
# Builds the reference probe (oracle/ref_probe.cpp + the reference's own
# src/fcwt/fcwt.cpp, compiled where it lies) into oracle/_ref/.
# Test infrastructure only.  Needs /root/reference; exits 0 quietly without it.
#
# Two variants, matching the two code paths of daughter_wavelet_multiplication:
#   ref_probe_avx     -O3 -mavx -fopenmp  (setup.py:45, CMakeLists.txt:56,141-157)
#   ref_probe_scalar  -O3 -fopenmp        (the non-AVX branch, fcwt.cpp:276-307)
# FFTW is not linked: --gc-sections drops every function that calls it, and the
# probe never calls them.  Only the header libs/fftw3.h (a source file) is used.
set -e
REF=${FCWT_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -f "$REF/src/fcwt/fcwt.cpp" ]; then
  echo "build_ref: reference sources not present; skipping" >&2
  exit 0
fi
mkdir -p "$OUT"
# fcwt.cpp uses assert/memset/memcpy without including <cassert>/<cstring>
# (it relies on transitive includes of older toolchains): force-include them.
# -DNDEBUG as in the shipped builds (asserts compiled out).
COMMON="-DNDEBUG -include cassert -include cstring -std=c++17 -fopenmp -ffunction-sections -fdata-sections -Wl,--gc-sections -iquote $REF/libs -I$REF/src/fcwt -w"
g++ -O3 -mavx $COMMON "$REF/src/fcwt/fcwt.cpp" "$HERE/ref_probe.cpp" -o "$OUT/ref_probe_avx"
g++ -O3 $COMMON "$REF/src/fcwt/fcwt.cpp" "$HERE/ref_probe.cpp" -o "$OUT/ref_probe_scalar"
