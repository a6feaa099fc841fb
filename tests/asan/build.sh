#!/bin/bash
# Build tests/asan/sar_asan_driver: the library's sources and the driver, with
# AddressSanitizer on the HOST code only (device code as in the product build).
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
B=$HERE/_build
mkdir -p "$B"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
FLAGS="-O1 -g -std=c++17 --offload-arch=gfx950 -fPIC -ffp-contract=off -munsafe-fp-atomics -Xclang -target-feature -Xclang -packed-fp32-ops -Wno-unused-command-line-argument -I $ROOT/include"
SAN="-Xarch_host -fsanitize=address -Xarch_host -fno-omit-frame-pointer"
objs=""
for s in sar_kernels sar_plan conv_kernels resample_kernels scene_kernels fft_kernels; do
  $HIPCC $FLAGS $SAN -c "$ROOT/sarsolver_amd/csrc/$s.hip" -o "$B/$s.o" 2> "$B/$s.log" &
  objs="$objs $B/$s.o"
done
$HIPCC $FLAGS $SAN -c "$HERE/sar_asan_driver.cpp" -x hip -o "$B/driver.o" 2> "$B/driver.log" &
wait
$HIPCC --offload-arch=gfx950 $SAN $objs "$B/driver.o" -o "$HERE/sar_asan_driver" \
  -L/opt/rocm/lib -lrocfft -Wl,-rpath,/opt/rocm/lib
echo "$HERE/sar_asan_driver"
