#!/bin/sh
# Build the reference atkdsp C library from its own sources into oracle/_ref/,
# for use as a checker only (golden vectors, the CPU baseline timing).
#
# Compiles the thirteen reference files the path and its next rows need, in place under
# /root/reference, with the flags of the reference's build.sh (C99, -O3, AVX2,
# FMA, OpenMP, never fast-math). Nothing is copied into this repository; the
# output (oracle/_ref/atkdsp_ref.so) is git-ignored. Exits 0 without building
# when the reference tree is absent.
set -e
REF=${ATKDSP_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -f "$REF/include/atkdsp.h" ]; then
    echo "build_ref: no reference tree at $REF; skipped" >&2
    exit 0
fi
mkdir -p "$OUT"
cd "$REF"
gcc -std=c99 -O3 -fno-math-errno -mavx2 -mfma -fopenmp -fPIC -shared \
    -fvisibility=hidden -DATKDSP_BUILD=1 -Iinclude -Ivendor/pocketfft \
    src/atkdsp.c src/convert.c src/channelizer.c src/fft.c src/display.c \
    src/detect.c src/design.c src/resample.c src/nco.c src/ring.c src/cfar.c src/stat.c vendor/pocketfft/pocketfft.c \
    -o "$OUT/atkdsp_ref.so.tmp" -lm -Wl,--no-undefined 2>"$OUT/build.log"
mv "$OUT/atkdsp_ref.so.tmp" "$OUT/atkdsp_ref.so"
echo "build_ref: $OUT/atkdsp_ref.so" >&2

# The reference's own ring stress test (tests/test_ring_stress.c), compiled
# unmodified with the reference header, twice: against the reference library
# and against this repository's drop-in (atkdsp_amd/atkdsp.so, when built).
gcc -std=c99 -D_POSIX_C_SOURCE=200809L -O2 -Iinclude tests/test_ring_stress.c -o "$OUT/ring_stress_ref" \
    "$OUT/atkdsp_ref.so" -Wl,-rpath,"$OUT" -lpthread 2>>"$OUT/build.log"
OURS="$HERE/../atkdsp_amd/atkdsp.so"
if [ -f "$OURS" ]; then
    OURS=$(cd "$(dirname "$OURS")" && pwd)/atkdsp.so
    gcc -std=c99 -D_POSIX_C_SOURCE=200809L -O2 -Iinclude tests/test_ring_stress.c -o "$OUT/ring_stress_ours" \
        "$OURS" -Wl,-rpath,"$(dirname "$OURS")" -lpthread 2>>"$OUT/build.log"
fi
echo "build_ref: $OUT/ring_stress_ref, ring_stress_ours" >&2
