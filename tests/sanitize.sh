#!/usr/bin/env bash
# CPU test suite under AddressSanitizer + UndefinedBehaviorSanitizer.
#
# Builds the host mesh builder, the command line and the oracle with
# -fsanitize=address,undefined (g++, host code only; no GPU is used), then runs
# `pytest -m "not gpu"` with those builds loaded in place of the regular ones and
# the sanitized command line on its argument-parsing paths.  Any report aborts
# the run (-fno-sanitize-recover=all, halt_on_error).
#
#   bash tests/sanitize.sh [extra pytest args]
set -euo pipefail
ROOT="$(cd "$(dirname "${BASH_SOURCE[0]}")/.." && pwd)"
make -s -C "$ROOT/fastercap_amd/csrc" sanitize
make -s -C "$ROOT/oracle" sanitize

SAN_DIR="$ROOT/fastercap_amd/lib/sanitize"
export FASTERCAP_HOST_LIB="$SAN_DIR/libfastercap_host.so"
export FASTERCAP_ORACLE_LIB="$ROOT/oracle/_build/sanitize/liboracle.so"
# the Python interpreter is not instrumented: the ASan runtime is preloaded so it
# comes first, and leak checking is off (the interpreter keeps its objects alive)
ASAN_RT="$(g++ -print-file-name=libasan.so)"
export ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=1:detect_odr_violation=0"
export UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1"

# command line: help, parse errors and a missing input (no device is opened)
CLI="$SAN_DIR/fastercap"
"$CLI" -b? > /dev/null
"$CLI" -bv > /dev/null
for bad in "" "-g x.lst" "-psx x.lst" "-mabc x.lst" "a.lst b.lst"; do
    set +e
    # shellcheck disable=SC2086
    "$CLI" $bad > /dev/null 2>&1
    rc=$?
    set -e
    if [ "$rc" -ne 64 ]; then
        echo "sanitized command line: '$bad' returned $rc, expected 64" >&2
        exit 1
    fi
done
echo "sanitized command line: parse paths clean"

cd "$ROOT"
LD_PRELOAD="$ASAN_RT" python -m pytest tests -m "not gpu" -x -q -p no:cacheprovider "$@"
