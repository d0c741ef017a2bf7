#!/bin/bash
# Builds libgbr.so (HIP kernels + host library + C ABI) for gfx950, in-tree.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(cd "$HERE/.." && pwd)"
OUT="$HERE/habitat_sim_amd/_lib"
mkdir -p "$OUT"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
FLAGS=(--offload-arch=gfx950 -O3 -fPIC -std=c++17 -ffp-contract=off
       -Xclang -target-feature -Xclang -packed-fp32-ops
       -I"$ROOT/include" -I"$HERE/csrc" -Wall -Wno-unused-function)
"$HIPCC" "${FLAGS[@]}" -c "$HERE/csrc/gbr_kernels.hip" -o "$OUT/gbr_kernels.o"
"$HIPCC" "${FLAGS[@]}" -c "$HERE/csrc/gbr_mlp.hip" -o "$OUT/gbr_mlp.o"
"$HIPCC" "${FLAGS[@]}" -c "$HERE/csrc/gbr_renderer.cpp" -o "$OUT/gbr_renderer.o"
"$HIPCC" "${FLAGS[@]}" -c "$HERE/csrc/gbr_vis.cpp" -o "$OUT/gbr_vis.o"
"$HIPCC" "${FLAGS[@]}" -c "$HERE/csrc/gbr_import.cpp" -o "$OUT/gbr_import.o"
"$HIPCC" "${FLAGS[@]}" -c "$HERE/csrc/gbr_noise.hip" -o "$OUT/gbr_noise.o"
"$HIPCC" --offload-arch=gfx950 -shared -fPIC -o "$OUT/libgbr.so.tmp" "$OUT/gbr_kernels.o" \
  "$OUT/gbr_mlp.o" "$OUT/gbr_renderer.o" "$OUT/gbr_vis.o" "$OUT/gbr_import.o" "$OUT/gbr_noise.o" -lz
mv "$OUT/libgbr.so.tmp" "$OUT/libgbr.so"
rm -f "$OUT/gbr_kernels.o" "$OUT/gbr_mlp.o" "$OUT/gbr_renderer.o" "$OUT/gbr_vis.o" "$OUT/gbr_import.o" \
  "$OUT/gbr_noise.o"
echo "built $OUT/libgbr.so"
