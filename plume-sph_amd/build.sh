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
"$HIPCC" "${FLAGS[@]}" -c "$HERE/csrc/psph_domain.cpp" -o "$HERE/_build/psph_domain.o"
"$HIPCC" "${FLAGS[@]}" -c "$HERE/csrc/psph_device.hip" -o "$HERE/_build/psph_device.o"
"$HIPCC" -shared --offload-arch=gfx950 "$HERE/_build/psph_host.o" "$HERE/_build/psph_domain.o" \
    "$HERE/_build/psph_device.o" -o "$OUT.tmp"
mv "$OUT.tmp" "$OUT"
echo "built $OUT"
# Output library (H5Part / CSV): host code against the image's HDF5 and libpsph.so.
# The run path lists the system library directory before HDF5's, so that the
# system libstdc++ (which ROCm needs) is found before the older one beside HDF5.
H5="${HDF5_DIR:-/opt/conda}"
SYSLIB="$(dirname "$(g++ -print-file-name=libstdc++.so.6)")"
SYSLIB="$(cd "$SYSLIB" && pwd -P)"
IO="$HERE/psph/libpsph_io.so"
g++ -O2 -std=c++17 -fPIC -shared -Wall "$HERE/csrc/psph_io.cpp" -I"$H5/include" \
    -L"$H5/lib" -lhdf5 -L"$HERE/psph" -lpsph \
    -Wl,-rpath,'$ORIGIN' -Wl,-rpath,"$SYSLIB" -Wl,-rpath,"$H5/lib" -o "$IO.tmp"
mv "$IO.tmp" "$IO"
echo "built $IO"
