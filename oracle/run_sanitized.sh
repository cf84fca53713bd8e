#!/bin/bash
# Runs the CPU test suite with the oracle built under AddressSanitizer and UndefinedBehaviorSanitizer
# (oracle/Makefile target `sanitize`).  The ASan runtime is preloaded into the Python interpreter
# that loads the library through ctypes; leak checking is off (the interpreter's own allocations
# outlive it), every other finding aborts the run.  Extra arguments go to pytest.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
make -s -C "$HERE" sanitize
ASAN_RT="$(gcc -print-file-name=libasan.so)"
UBSAN_RT="$(gcc -print-file-name=libubsan.so)"
cd "$HERE/.."
CUTCELL_ORACLE_LIB="$HERE/liborc_san.so" \
ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1" \
UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1" \
LD_PRELOAD="$ASAN_RT $UBSAN_RT${LD_PRELOAD:+ $LD_PRELOAD}" \
python3 -m pytest tests -q -m "not gpu" -p no:cacheprovider "$@"
