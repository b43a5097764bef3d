#!/bin/bash
# Builds libpa_hip.so for gfx950 in-tree (pyatmosphere_amd/libpa_hip.so).
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
out="$here/../libpa_hip.so"
obj="$here/_build"
mkdir -p "$obj"
FLAGS=(--offload-arch=gfx950 -O3 -fPIC -std=c++17 -Wall -Wno-unused-function
       -Xclang -target-feature -Xclang -packed-fp32-ops -I"$here")
pids=()
for src in pa_step pa_screen pa_reduce pa_api pa_mlp pa_calib pa_channel; do
  [ -f "$here/$src.hip" ] || continue
  hipcc "${FLAGS[@]}" -c "$here/$src.hip" -o "$obj/$src.o" 2> >(grep -v "packed-fp32-ops' is not a recognized" >&2) &
  pids+=($!)
done
for p in "${pids[@]}"; do wait "$p"; done
hipcc --offload-arch=gfx950 -shared -fPIC "$obj"/*.o -o "$out.tmp"
mv "$out.tmp" "$out"
echo "built $out"
