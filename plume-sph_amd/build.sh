#!/bin/bash
# Build the MI355X (gfx950) library libpsph.so in-tree.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/psph/libpsph.so"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
FLAGS=(-O3 -std=c++17 -fPIC --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops
       -ffp-contract=fast -Wall -Wno-unused-function -Wno-unused-variable -Wno-unused-result -Wno-unused-value)
mkdir -p "$HERE/_build"
"$HIPCC" "${FLAGS[@]}" -c "$HERE/csrc/psph_host.cpp" -o "$HERE/_build/psph_host.o"
"$HIPCC" "${FLAGS[@]}" -c "$HERE/csrc/psph_device.hip" -o "$HERE/_build/psph_device.o"
"$HIPCC" -shared --offload-arch=gfx950 "$HERE/_build/psph_host.o" "$HERE/_build/psph_device.o" -o "$OUT.tmp"
mv "$OUT.tmp" "$OUT"
echo "built $OUT"
