#!/bin/sh
# Builds the drop-in C++ test against the built library and the oracle.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
g++ -std=c++20 -O1 -Wall -I "$ROOT/include" "$HERE/dropin_test.cpp" -o "$HERE/dropin_test" \
    -L "$ROOT/kascade_amd" -lkascade_amd -L "$ROOT/oracle" -loracle \
    -Wl,-rpath,'$ORIGIN/../../kascade_amd' -Wl,-rpath,'$ORIGIN/../../oracle' -lpthread
