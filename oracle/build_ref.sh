#!/usr/bin/env bash
# Compile the reference solvers from their own sources into oracle/_ref/.
#
# Test infrastructure only: the binaries are the checker used to write and
# re-validate golden vectors (tests/golden/make_golden.py) and an optional CPU
# baseline. Nothing in the product links or runs them. The sources are read
# where they lie; nothing is copied into this repository.
#
# The reference compiles with the standard library alone. Its headers rely on
# a transitive <cstddef> for size_t under MinGW, so the one addition is the
# compiler flag `-include cstddef`; no source is edited.
set -euo pipefail
REF=${CR_REFERENCE_DIR:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF/solvers" ]; then
    echo "build_ref: no reference tree at $REF; skipping" >&2
    exit 0
fi
mkdir -p "$OUT"
CXX=${CXX:-g++}
for s in k_cops_visibility k_cops_visibility_v2; do
    "$CXX" -std=c++17 -include cstddef -O3 -I"$REF/lib/include" \
        "$REF/app/main.cpp" "$REF"/lib/src/*.cpp \
        -DALGORITHM_INCLUDE="\"$REF/solvers/$s.cpp\"" -o "$OUT/$s"
done
# The reference's multithreaded level-synchronous solver (stay-move rules).
"$CXX" -std=c++17 -include cstddef -O3 -pthread -I"$REF/legacy/include" \
    "$REF/legacy/src/k_cops_4.cpp" "$REF/legacy/src/Graph.cpp" \
    "$REF/legacy/src/AdjacencyList.cpp" "$REF/legacy/src/fileio.cpp" -o "$OUT/k_cops_4"
# The reference's own supervisor (app/main.cpp) with this package's plug-in
# (copsandrobbers_amd/csrc/plugin_solver.cpp) as the solver translation unit:
# the drop-in as a maintainer of the reference would build it, linked against
# libcrsolve.so. Built only when the library is already there.
REPO=$(cd "$HERE/.." && pwd)
if [ -f "$REPO/copsandrobbers_amd/libcrsolve.so" ]; then
    "$CXX" -std=c++17 -include cstddef -O3 -I"$REF/lib/include" -I"$REPO/include" \
        "$REF/app/main.cpp" "$REF"/lib/src/*.cpp \
        -DALGORITHM_INCLUDE="\"$REPO/copsandrobbers_amd/csrc/plugin_solver.cpp\"" \
        -L"$REPO/copsandrobbers_amd" -lcrsolve -Wl,-rpath,'$ORIGIN/../../copsandrobbers_amd' \
        -o "$OUT/k_cops_visibility_dropin"
fi
echo "build_ref: built $(ls "$OUT" | tr '\n' ' ')"
