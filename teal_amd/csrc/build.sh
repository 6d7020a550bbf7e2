#!/bin/bash
# Build libteal_amd.so for gfx950 in place (teal_amd/libteal_amd.so).
#   FP64 kernels: -ffp-contract=off (bit-compatible with the reference)
#   FP32 kernels: contraction on, FP32 division and square root to within a few ulp (the
#                 hardware reciprocal and square-root estimates) instead of correctly rounded,
#                 FP32 denormals flushed to zero (FP64 arithmetic keeps IEEE denormals)
# Packed FP32 instructions are disabled (-packed-fp32-ops) on every HIP line.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
OUT="$HERE/../libteal_amd.so"
OBJ="$HERE/../_build"
mkdir -p "$OBJ"
ARCH="${TEAL_ARCH:-gfx950}"
JOBS="${TEAL_JOBS:-8}"
HIPFLAGS=(-O3 -std=c++17 -fPIC --offload-arch="$ARCH" -I"$ROOT/include" -I"$HERE"
          -Xclang -target-feature -Xclang -packed-fp32-ops -Wno-unused-result)
CFLAGS=(-std=c11 -O2 -fPIC -ffp-contract=off -Wall -Wextra -Wno-unused-parameter -I"$ROOT/include")

pids=()
build_hip() {  # src flags...
    local src="$1"; shift
    /opt/rocm/bin/hipcc "${HIPFLAGS[@]}" "$@" -c "$HERE/$src" -o "$OBJ/${src%.*}.o"
}
build_hip device.hip -ffp-contract=off & pids+=($!)
build_hip mixed.hip -ffp-contract=fast -fno-hip-fp32-correctly-rounded-divide-sqrt \
    -fgpu-flush-denormals-to-zero & pids+=($!)
build_hip krylov.hip -ffp-contract=off & pids+=($!)
build_hip plugins.hip -ffp-contract=off & pids+=($!)
build_hip diagnostics.hip -ffp-contract=off & pids+=($!)
# VTKHDF output and restart: the image's serial HDF5 headers (/opt/conda); the library itself
# is loaded at run time (io.c)
H5DIR="${TEAL_HDF5:-/opt/conda}"
gcc "${CFLAGS[@]}" -I"$H5DIR/include" -c "$HERE/io.c" -o "$OBJ/io.o" & pids+=($!)
for src in core.c sync.c mesh_create.c mesh_generate.c mesh_read.c mesh_util.c mesh_union.c equations.c simulation.c; do
    if [ -f "$HERE/$src" ]; then
        gcc "${CFLAGS[@]}" -c "$HERE/$src" -o "$OBJ/${src%.*}.o" & pids+=($!)
    fi
done
rc=0
for p in "${pids[@]}"; do wait "$p" || rc=1; done
[ $rc -eq 0 ] || { echo "build failed" >&2; exit 1; }
/opt/rocm/bin/hipcc -shared -fPIC --offload-arch="$ARCH" -o "$OUT.tmp" "$OBJ"/*.o -lm -ldl -Wl,--no-undefined
mv "$OUT.tmp" "$OUT"
echo "built $OUT"
