#!/bin/bash
# Builds the Fortran host (modules with the reference's interfaces + the
# channel_FD driver) against libcfhip.so with AMD flang.  When the
# reference sources are mounted, the reference's own main program
# (main.f90) is compiled from where it lies and linked against the same
# modules (lib/channel_fd_ref): the drop-in check of the module interfaces.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
FC=${FC:-/opt/rocm/lib/llvm/bin/flang}
REF=${REF:-/root/reference}
OUT="$HERE/../lib"
OBJ="$HERE/../_build/fortran"
mkdir -p "$OBJ" "$OUT"
cd "$OBJ"
MODS=(cf_hip global hot_path io_driver driver_modules)
for f in "${MODS[@]}" main; do
  "$FC" -O2 -fPIC -c "$HERE/$f.f90" -o "$OBJ/$f.o"
done
objs=()
for f in "${MODS[@]}"; do objs+=("$OBJ/$f.o"); done
"$FC" -O2 -o "$OUT/channel_fd" "${objs[@]}" "$OBJ"/main.o \
  -L"$OUT" -lcfhip -Wl,-rpath,'$ORIGIN' -Wl,-rpath,/opt/rocm/lib
# module library for hosts that bring their own main program
"$FC" -shared -fPIC -o "$OUT/libcf_fortran.so.tmp" "${objs[@]}" -L"$OUT" -lcfhip -Wl,-rpath,'$ORIGIN'
mv "$OUT/libcf_fortran.so.tmp" "$OUT/libcf_fortran.so"
echo "built $OUT/channel_fd"
if [ -f "$REF/main.f90" ]; then
  mkdir -p "$OBJ/ref"
  "$FC" -O2 -c "$REF/main.f90" -I"$OBJ" -module-dir "$OBJ/ref" -o "$OBJ/ref/main.o"
  "$FC" -O2 -o "$OUT/channel_fd_ref" "${objs[@]}" "$OBJ/ref/main.o" \
    -L"$OUT" -lcfhip -Wl,-rpath,'$ORIGIN' -Wl,-rpath,/opt/rocm/lib
  echo "built $OUT/channel_fd_ref (reference main.f90 over these modules)"
fi
