#!/bin/sh
# Builds yang-mills_amd/libym_amd.so for gfx950.
# -ffp-contract=off: every fused multiply-add in the kernels is an explicit
#   __builtin_fma/__builtin_fmaf, so results are reproducible by a CPU replay.
# -packed-fp32-ops disabled: no v_pk_*_f32 instructions are generated.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
SRC="$HERE/csrc/ym_su3.hip $HERE/csrc/ym_su2.hip $HERE/csrc/ym_obs.hip $HERE/csrc/ym_capi.hip $HERE/csrc/ym_ensemble.hip $HERE/csrc/ym_learned.hip $HERE/csrc/ym_host_stats.cpp $HERE/csrc/ym_host_fit.cpp $HERE/csrc/ym_codec.cpp"
$HIPCC --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared \
  -ffp-contract=off -fno-gpu-flush-denormals-to-zero \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  -Wall -Wno-unused-function -Wno-unused-variable \
  -o "$HERE/libym_amd.so.tmp" $SRC
mv "$HERE/libym_amd.so.tmp" "$HERE/libym_amd.so"
