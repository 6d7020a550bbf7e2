#!/bin/bash
# Build the C drivers of this directory against include/ and teal_amd/libteal_amd.so
# (tests/drivers/bin/<name>); called by __graft_entry__.build().
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
mkdir -p "$HERE/bin"
for src in "$HERE"/*.c; do
    name="$(basename "${src%.c}")"
    gcc -std=c11 -O2 -Wall -Wextra -Wno-unused-parameter -I"$ROOT/include" "$src" \
        -o "$HERE/bin/$name" -L"$ROOT/teal_amd" -lteal_amd -Wl,-rpath,'$ORIGIN/../../../teal_amd' -lm
done
