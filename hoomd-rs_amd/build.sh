#!/bin/bash
# Build libhmc.so (gfx950) in-tree: one object per translation unit, compiled
# in parallel, then linked.  Usage: hoomd-rs_amd/build.sh
set -euo pipefail
cd "$(dirname "$0")"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
JOBS=${JOBS:-8}
FLAGS=(-O3 -std=c++17 --offload-arch=gfx950 -fPIC -ffp-contract=off -fno-fast-math
       -Xclang -target-feature -Xclang -packed-fp32-ops
       -Wall -Wno-unused-function -Wno-unused-parameter -Wno-unused-result)
SRCS=(csrc/k_build.hip csrc/k_hard.hip csrc/k_persist.hip csrc/k_persist_lj.hip csrc/k_lj.hip csrc/k_convex.hip
      csrc/k_measure.hip csrc/k_misc.hip csrc/api_ctx.hip csrc/api_batch.hip csrc/api_measure.hip
      csrc/api_place.hip csrc/api_gsd.hip csrc/k_tune.hip)
mkdir -p _build
pids=()
objs=()
for src in "${SRCS[@]}"; do
  obj="_build/$(basename "${src%.hip}").o"
  objs+=("$obj")
  "$HIPCC" "${FLAGS[@]}" -c -o "$obj" "$src" &
  pids+=($!)
  if (( ${#pids[@]} >= JOBS )); then
    wait "${pids[0]}"
    pids=("${pids[@]:1}")
  fi
done
for p in "${pids[@]}"; do wait "$p"; done
"$HIPCC" --offload-arch=gfx950 -fPIC -shared -o libhmc.so.tmp "${objs[@]}"
mv libhmc.so.tmp libhmc.so
