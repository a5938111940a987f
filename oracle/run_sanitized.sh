#!/bin/bash
# The oracle's CPU tests with the oracle built under AddressSanitizer and
# UndefinedBehaviorSanitizer: the sanitized extension (make sanitize) is put
# first on the module path and the sanitizer runtimes are preloaded into the
# Python process (leak detection off: the interpreter itself keeps memory to
# the end).  Usage, from the repository root: bash oracle/run_sanitized.sh
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
root="$(dirname "$here")"
make -s -C "$here" sanitize
asan="$(gcc -print-file-name=libasan.so)"
ubsan="$(gcc -print-file-name=libubsan.so)"
cd "$root"
PYTHONPATH="$here/_build/asan:$root/tests:$root" \
LD_PRELOAD="$asan $ubsan" \
ASAN_OPTIONS=detect_leaks=0:halt_on_error=1 \
UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1 \
  python3 -m pytest -q -p no:cacheprovider tests/test_oracle.py tests/test_golden.py -m "not gpu" "$@"
