#!/bin/bash
# Builds the library with AddressSanitizer on its host code only (every
# -fsanitize applies to the host side; device code is compiled as usual) and
# the driver tests/cpp/asan_driver.cpp against it, into tests/cpp/_asan/.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
OUT="$HERE/_asan"
SRC="$ROOT/kascade_amd/csrc"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
mkdir -p "$OUT"
FLAGS="-O1 -g -fPIC -std=c++20 --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops -I $ROOT/include -Xarch_host -fsanitize=address -Xarch_host -fno-omit-frame-pointer"
objs=""
for f in api.cpp rank.hip gen.hip util.hip euler.hip segmented.hip batch.hip server.cpp probe.hip listutil.hip graph.hip \
         noalloc.hip contract.hip; do
  "$HIPCC" -x hip $FLAGS -c "$SRC/$f" -o "$OUT/$f.o" &
  objs="$objs $OUT/$f.o"
done
wait
"$HIPCC" --offload-arch=gfx950 -shared -fPIC -fsanitize=address -fno-gpu-sanitize -o "$OUT/libkascade_amd.so" $objs
"$HIPCC" -O1 -g -std=c++20 -I "$ROOT/include" -Xarch_host -fsanitize=address "$HERE/asan_driver.cpp" \
    -o "$OUT/asan_driver" -L "$OUT" -lkascade_amd -Wl,-rpath,'$ORIGIN' -fno-gpu-sanitize
