#!/bin/bash
# Builds libcfhip.so (HIP kernels for gfx950 + C-ABI) in-tree.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
SRC="$HERE/csrc"
OUT="$HERE/lib"
OBJ="$HERE/_build"
mkdir -p "$OUT" "$OBJ"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
ARCH=${CF_ARCH:-gfx950}
FLAGS=(-O3 -fPIC -std=c++17 --offload-arch="$ARCH" -Xclang -target-feature -Xclang -packed-fp32-ops
       -munsafe-fp-atomics -ffp-contract=fast -Wall -Wno-unused-function -Wno-unused-variable -Wno-unused-value -Wno-unused-result)
pids=()
for f in cf_stencil cf_fft cf_dsm cf_eqwm cf_mlp cf_api cf_serve; do
  src="$SRC/$f.hip"; [ -f "$src" ] || src="$SRC/$f.cpp"
  [ -f "$src" ] || continue
  "$HIPCC" "${FLAGS[@]}" -x hip -c "$src" -o "$OBJ/$f.o" &
  pids+=($!)
done
for p in "${pids[@]}"; do wait "$p"; done
objs=()
for f in cf_stencil cf_fft cf_dsm cf_eqwm cf_mlp cf_api cf_serve; do
  [ -f "$OBJ/$f.o" ] && objs+=("$OBJ/$f.o")
done
"$HIPCC" --offload-arch="$ARCH" -shared -fPIC -o "$OUT/libcfhip.so.tmp" "${objs[@]}"
mv "$OUT/libcfhip.so.tmp" "$OUT/libcfhip.so"
echo "built $OUT/libcfhip.so"
