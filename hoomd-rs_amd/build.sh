#!/bin/bash
# Build libhmc.so (gfx950) in-tree.  Usage: hoomd-rs_amd/build.sh
set -euo pipefail
cd "$(dirname "$0")"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
$HIPCC -O3 -std=c++17 --offload-arch=gfx950 -fPIC -shared -ffp-contract=off -fno-fast-math \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  -Wall -Wno-unused-function -Wno-unused-parameter -Wno-unused-result \
  -o libhmc.so.tmp csrc/hmc_sweep.hip
mv libhmc.so.tmp libhmc.so
