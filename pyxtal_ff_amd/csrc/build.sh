#!/bin/bash
# Build libpxff.so for gfx950 (hipcc).  Usage: build.sh [outdir]
# Translation units compile in parallel; the link waits for all of them.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="${1:-$HERE/../_lib}"
mkdir -p "$OUT" "$HERE/_build"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
# -packed-fp32-ops: hipcc must not emit packed FP32 instructions on this pool
FLAGS=(-O3 -std=c++17 -fPIC --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops -Wall -Wno-unused-result)
OBJS=()
PIDS=()
for src in pxff_api.cpp so4_tables.cpp nlist.hip so4_z.hip so4_kernels.hip so3_kernels.hip reduce.hip mlp.hip; do
  [ -f "$HERE/$src" ] || continue
  obj="$HERE/_build/${src%.*}.o"
  # PXFF_INCREMENTAL=1: keep objects newer than their source and every header
  if [ "${PXFF_INCREMENTAL:-0}" = 1 ] && [ -f "$obj" ] && [ "$obj" -nt "$HERE/$src" ]; then
    stale=0
    for h in "$HERE"/*.h "$HERE/../../include/pxff.h"; do [ "$h" -nt "$obj" ] && stale=1; done
    if [ $stale = 0 ]; then OBJS+=("$obj"); continue; fi
  fi
  EXTRA=()
  # the network kernel's FP64 MFMA accumulators stay in VGPRs (no copies
  # through the accumulation registers around every GEMM step)
  [ "$src" = mlp.hip ] && EXTRA=(-mllvm -amdgpu-mfma-vgpr-form)
  if [[ "$src" == *.hip ]]; then
    "$HIPCC" "${FLAGS[@]}" "${EXTRA[@]}" -c "$HERE/$src" -o "$obj" &
  else
    "$HIPCC" "${FLAGS[@]}" -x hip -c "$HERE/$src" -o "$obj" &
  fi
  PIDS+=($!)
  OBJS+=("$obj")
done
for p in "${PIDS[@]}"; do wait "$p"; done
"$HIPCC" -shared --offload-arch=gfx950 -o "$OUT/libpxff.so.tmp" "${OBJS[@]}"
mv "$OUT/libpxff.so.tmp" "$OUT/libpxff.so"
echo "built $OUT/libpxff.so"
