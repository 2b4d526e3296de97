#!/bin/bash
# Runs the CPU test suite with AddressSanitizer + UndefinedBehaviorSanitizer
# builds of the oracle and of the untrusted-file parsers (CPU only; GPU code
# is never built with sanitizers). Any sanitizer report aborts the run.
#   bash tests/sanitize/run.sh [extra pytest args]
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
make -s -C "$HERE" -j4
# the ASan runtime first; libstdc++ loaded with it so that its exception
# interceptors find the real __cxa_throw (python itself is not C++)
ASAN_RT="$(g++ -print-file-name=libasan.so):$(g++ -print-file-name=libstdc++.so)"
cd "$ROOT"
export HR_ORACLE_LIB="$HERE/_build/liboracle_hr.so"
export HR_SAN_HOST_LIB="$HERE/_build/libhr_host.so"
export ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1:detect_odr_violation=0"
export UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1"
export PYTHONPATH="$HERE${PYTHONPATH:+:$PYTHONPATH}"
LD_PRELOAD="$ASAN_RT${LD_PRELOAD:+:$LD_PRELOAD}" python -m pytest tests -m "not gpu" -q -p plugin \
    --ignore=tests/sanitize "$@"
