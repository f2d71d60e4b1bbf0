#!/bin/sh
# Builds the reference's own table1 C kernels into oracle/_ref/ for the oracle.
#
# Source: mrrt/nufft/_extensions/c/nufft_table.c (+ nufft_table.template.c,
# nufft_table.h, nufft_table.template.h, templating.h), compiled where it lies
# under /root/reference with the reference's flags: -ffast-math (setup.py:110)
# and -fopenmp (setup.py:36-50), plus the CPython default -O2 -fwrapv.
# Nothing is copied; the output is a shared library in oracle/_ref/ only.
# Without /root/reference (e.g. on a machine that only runs tests) this is a
# no-op and the previously built library, if any, is kept.
set -e
SRC=/root/reference/mrrt/nufft/_extensions/c
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -f "$SRC/nufft_table.c" ]; then
    echo "build_ref: reference sources not present; skipping"
    exit 0
fi
mkdir -p "$OUT"
gcc -O2 -fwrapv -fPIC -shared -ffast-math -fopenmp -I"$SRC" \
    "$SRC/nufft_table.c" -o "$OUT/libnufft_table_ref.so.tmp"
mv "$OUT/libnufft_table_ref.so.tmp" "$OUT/libnufft_table_ref.so"
echo "build_ref: built $OUT/libnufft_table_ref.so"
