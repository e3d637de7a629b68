#!/bin/bash
# Builds tests/host_asan/vx_host_check: the library's sources and the check
# driver with AddressSanitizer and UBSan in the host code only (-Xarch_host
# before each -fsanitize=), device code for gfx950 as in the library build.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(dirname "$(dirname "$HERE")")
SRC=$ROOT/wgpu-voxel-engine_amd/csrc
OUT=$HERE/_build
mkdir -p "$OUT"
SAN="-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -fno-sanitize-recover=undefined"
FLAGS="-O2 -g -std=c++17 -fno-omit-frame-pointer --offload-arch=gfx950 -ffp-contract=off -Xclang -target-feature -Xclang -packed-fp32-ops -I $ROOT/include -I $SRC"
for s in vx_kernels.hip vx_api.cpp vx_batch.cpp vx_server.cpp vx_world.cpp; do
    hipcc $FLAGS $SAN -x hip -c "$SRC/$s" -o "$OUT/${s%.*}.o" &
done
hipcc $FLAGS $SAN -x hip -c "$HERE/vx_host_check.cpp" -o "$OUT/vx_host_check.o" &
wait
hipcc --offload-arch=gfx950 $SAN "$OUT"/*.o -o "$HERE/vx_host_check"
echo "built $HERE/vx_host_check"
