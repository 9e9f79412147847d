#!/bin/sh
# Builds yang-mills_amd/libym_amd.so for gfx950: every source is compiled on
# its own (in parallel) into _build/, then the objects are linked.
# -ffp-contract=off: every fused multiply-add in the kernels is an explicit
#   __builtin_fma/__builtin_fmaf, so results are reproducible by a CPU replay.
# -packed-fp32-ops disabled: no v_pk_*_f32 instructions are generated.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
OUT="$HERE/_build"
mkdir -p "$OUT"
FLAGS="--offload-arch=gfx950 -O3 -std=c++17 -fPIC -ffp-contract=off -fno-gpu-flush-denormals-to-zero \
  -Xclang -target-feature -Xclang -packed-fp32-ops -Wall -Wno-unused-function -Wno-unused-variable"
SRCS="ym_su3.hip ym_su2.hip ym_obs.hip ym_capi.hip ym_ensemble.hip ym_learned.hip ym_host_stats.cpp ym_host_fit.cpp ym_codec.cpp"
PIDS=""
OBJS=""
for s in $SRCS; do
  o="$OUT/${s%.*}.o"
  OBJS="$OBJS $o"
  case "$s" in
    *.hip) $HIPCC $FLAGS -c "$HERE/csrc/$s" -o "$o" & ;;
    *) $HIPCC -O3 -std=c++17 -fPIC -ffp-contract=off -Wall -c "$HERE/csrc/$s" -o "$o" & ;;
  esac
  PIDS="$PIDS $!"
done
FAIL=0
for p in $PIDS; do
  wait "$p" || FAIL=1
done
[ "$FAIL" = 0 ] || { echo "build.sh: compilation failed" >&2; exit 1; }
$HIPCC --offload-arch=gfx950 -shared -fPIC -o "$HERE/libym_amd.so.tmp" $OBJS
mv "$HERE/libym_amd.so.tmp" "$HERE/libym_amd.so"
