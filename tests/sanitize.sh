#!/bin/bash
# The CPU half of the test suite against the host library and the oracle
# built with AddressSanitizer and UndefinedBehaviorSanitizer: the host
# Domain's set-up (node-to-corner lists, region runs, boundary masks), its
# driver loop, the VTK writer and the oracle's index arithmetic run
# instrumented.  The device library is not instrumented (GPU code).
#   bash tests/sanitize.sh [extra pytest arguments]
set -euo pipefail
ROOT="$(cd "$(dirname "$0")/.." && pwd)"
python3 -m lulesh_amd.build --sanitize > /dev/null
make -s -C "$ROOT/oracle" asan
ASAN_LIB="$(gcc -print-file-name=libasan.so)"
UBSAN_LIB="$(gcc -print-file-name=libubsan.so)"
cd "$ROOT"
# detect_leaks=0: the Python interpreter and the HIP runtime keep allocations
# alive at exit by design
LD_PRELOAD="$ASAN_LIB:$UBSAN_LIB" \
ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1" \
UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1" \
LULESH_LIB_DIR="$ROOT/lulesh_amd/lib_asan" \
ORACLE_SO_PATH="$ROOT/oracle/build_asan/liblulesh_oracle.so" \
python3 -m pytest tests -q -m "not gpu" -p no:cacheprovider "$@"
