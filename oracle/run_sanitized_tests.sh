#!/bin/bash
# Runs the CPU tests of the oracle against its AddressSanitizer +
# UndefinedBehaviorSanitizer build.  Python itself is not instrumented, so the
# ASan runtime is preloaded; leak checking is off (the interpreter's own
# allocations would be reported).  Exit status is pytest's.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(dirname "$HERE")"
make -s -C "$HERE" sanitize
ASAN_RT="$(gcc -print-file-name=libasan.so)"
UBSAN_RT="$(gcc -print-file-name=libubsan.so)"
cd "$ROOT"
LD_PRELOAD="$ASAN_RT:$UBSAN_RT" \
ASAN_OPTIONS=detect_leaks=0:abort_on_error=1:halt_on_error=1 \
UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1 \
SSG_ORACLE_LIB="$HERE/_san/libssg_oracle_san.so" \
    python -m pytest -q -p no:cacheprovider tests/test_oracle.py "$@"
