#!/bin/bash
# Builds the Fortran host (modules with the reference's interfaces + the
# channel_FD driver) against libcfhip.so with AMD flang.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
OUT="$HERE/../lib"
OBJ="$HERE/../_build/fortran"
mkdir -p "$OBJ" "$OUT"
cd "$OBJ"
for f in cf_hip global hot_path io_driver main; do
  "$FC" -O2 -fPIC -c "$HERE/$f.f90" -o "$OBJ/$f.o"
done
"$FC" -O2 -o "$OUT/channel_fd" "$OBJ"/cf_hip.o "$OBJ"/global.o "$OBJ"/hot_path.o "$OBJ"/io_driver.o "$OBJ"/main.o \
  -L"$OUT" -lcfhip -Wl,-rpath,'$ORIGIN' -Wl,-rpath,/opt/rocm/lib
# module library for hosts that bring their own main program
"$FC" -shared -fPIC -o "$OUT/libcf_fortran.so.tmp" "$OBJ"/cf_hip.o "$OBJ"/global.o "$OBJ"/hot_path.o \
  "$OBJ"/io_driver.o -L"$OUT" -lcfhip -Wl,-rpath,'$ORIGIN' 2>/dev/null || true
[ -f "$OUT/libcf_fortran.so.tmp" ] && mv "$OUT/libcf_fortran.so.tmp" "$OUT/libcf_fortran.so"
echo "built $OUT/channel_fd"
