#!/bin/bash
# Host-side AddressSanitizer build of the library and of the C ABI driver
# (device code is built normally: GPU ASan is not used).  Outputs go to
# tests/c_abi/_asan/.  Run:  ASAN_OPTIONS=detect_leaks=1 tests/c_abi/_asan/abi_driver
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
ROCM="${ROCM_PATH:-/opt/rocm}"
OUT="$HERE/_asan"
mkdir -p "$OUT"
HIPFLAGS=(--offload-arch=gfx950 -O1 -g -std=c++17 -fPIC -mcode-object-version=5
  -ffp-contract=off -Xclang -target-feature -Xclang -packed-fp32-ops
  -Wno-unused-command-line-argument
  -Xarch_host -fsanitize=address -Xarch_host -fno-omit-frame-pointer)
SRC="$ROOT/pytorch-lbm_amd/csrc"
"$ROCM/bin/hipcc" "${HIPFLAGS[@]}" -c -o "$OUT/lbm_api.o" "$SRC/lbm_api.hip"
for k in 0 1 2 3; do
  "$ROCM/bin/hipcc" "${HIPFLAGS[@]}" -DLBM_INST_PREC=$k -c -o "$OUT/lbm_inst_$k.o" "$SRC/lbm_inst.hip"
done
"$ROCM/bin/hipcc" "${HIPFLAGS[@]}" -shared -o "$OUT/liblbm_amd.so" "$OUT/lbm_api.o" \
  "$OUT"/lbm_inst_0.o "$OUT"/lbm_inst_1.o "$OUT"/lbm_inst_2.o "$OUT"/lbm_inst_3.o
"$ROCM/llvm/bin/clang++" -O1 -g -fsanitize=address -fno-omit-frame-pointer -std=c++17 \
  -D__HIP_PLATFORM_AMD__ -I"$ROCM/include" "$HERE/abi_driver.cpp" -o "$OUT/abi_driver" \
  -L"$OUT" -llbm_amd -L"$ROCM/lib" -lamdhip64 -Wl,-rpath,'$ORIGIN' -Wl,-rpath,"$ROCM/lib"
echo "built $OUT/abi_driver"
