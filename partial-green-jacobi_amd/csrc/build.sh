#!/bin/sh
# Builds libpgj.so for gfx950 from the HIP sources in this directory.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=${1:-$HERE/../pgj/libpgj.so}
FLAGS="--offload-arch=gfx950 -O3 -std=c++17 -fPIC -Xclang -target-feature -Xclang -packed-fp32-ops"
TMP=$(mktemp -d)
trap 'rm -rf "$TMP"' EXIT
PIDS=""
for src in common hex8 hex8_mfma fft fft_static resident oc pcg mixed select hybrid; do
  hipcc $FLAGS -c "$HERE/$src.hip" -o "$TMP/$src.o" &
  PIDS="$PIDS $!"
done
FAIL=0
for p in $PIDS; do wait "$p" || FAIL=1; done
if [ "$FAIL" -ne 0 ]; then echo "build.sh: compilation failed" >&2; exit 1; fi
hipcc --offload-arch=gfx950 -shared -o "$OUT.tmp" "$TMP"/common.o "$TMP"/hex8.o "$TMP"/hex8_mfma.o "$TMP"/fft.o "$TMP"/fft_static.o "$TMP"/resident.o "$TMP"/oc.o "$TMP"/pcg.o "$TMP"/mixed.o "$TMP"/select.o "$TMP"/hybrid.o
mv "$OUT.tmp" "$OUT"
