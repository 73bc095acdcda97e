#!/bin/bash
# Builds the host side of the library against the HIP stand-in (mockhip.cpp),
# with AddressSanitizer on the host code, plus the serving replay driver:
#   tests/standin/_build/libapocharmm_standin.so, libmockhip.so, serve_replay
# Host-only compilation (--cuda-host-only): no device code is built.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
OUT="$HERE/_build"
mkdir -p "$OUT"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
SAN="-fsanitize=address -fno-omit-frame-pointer"
g++ -std=c++17 -O1 -g -fPIC $SAN -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include -shared \
  "$HERE/mockhip.cpp" -o "$OUT/libmockhip.so"
objs=()
pids=()
for s in system.hip util.hip nblist.hip pair.hip bonded.hip pme.hip md.hip capi.cpp \
         charmm_files.cpp capi_charmm.cpp noalloc.hip; do
  "$HIPCC" --cuda-host-only -O1 -g -std=c++17 -fPIC -Xarch_host -fsanitize=address \
    -Xarch_host -fno-omit-frame-pointer -I"$ROOT/include" -x hip -c "$ROOT/apocharmm_amd/csrc/$s" \
    -o "$OUT/$s.o" &
  pids+=($!)
  objs+=("$OUT/$s.o")
done
for p in "${pids[@]}"; do wait "$p"; done
# the host objects name the device code's fat binaries; none exists here
nm "${objs[@]}" | awk '$1 == "U" && $2 ~ /^__hip_fatbin/ {print $2}' | sort -u |
  awk '{print "char " $1 "[64];"}' > "$OUT/fatbins.c"
gcc -c -fPIC "$OUT/fatbins.c" -o "$OUT/fatbins.o"
g++ -shared -fPIC $SAN -o "$OUT/libapocharmm_standin.so" "${objs[@]}" "$OUT/fatbins.o" \
  -L"$OUT" -lmockhip -Wl,-rpath,"$OUT"
g++ -std=c++17 -O1 -g $SAN -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include -I"$ROOT/include" \
  "$HERE/serve_replay.cpp" -o "$OUT/serve_replay" -L"$OUT" -lapocharmm_standin -lmockhip \
  -Wl,-rpath,"$OUT"
echo "built $OUT/serve_replay"
