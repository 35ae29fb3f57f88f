#!/bin/bash
# The CPU test suite against AddressSanitizer + UndefinedBehaviorSanitizer
# builds of the oracle (oracle/_san/liboracle.so) and of the native adjoint
# checkers (tests/native/_san/*.so).  Python itself is not instrumented, so the
# sanitizer runtimes are preloaded; leak checking is off (the interpreter keeps
# its allocations to the end).  Any report fails the run.
#
#   oracle/run_sanitized.sh [pytest args]     (default: the oracle and native-checker tests)
set -euo pipefail
cd "$(dirname "$0")/.."
make -s -C oracle sanitize
make -s -C tests/native sanitize
export LD_PRELOAD="$(gcc -print-file-name=libasan.so) $(gcc -print-file-name=libubsan.so)"
export ASAN_OPTIONS=detect_leaks=0:abort_on_error=1:halt_on_error=1
export UBSAN_OPTIONS=halt_on_error=1:print_stacktrace=1
export OMHD_SANITIZED=1
if [ $# -eq 0 ]; then
  set -- tests/test_oracle_golden.py tests/test_oracle_resistive3d.py tests/test_adjoint_cpu.py tests/test_flopcount.py \
    tests/test_brio_wu.py tests/test_recon2d.py tests/test_symmetry.py tests/test_invariants.py tests/test_snapshot.py \
    tests/test_diagnostics.py -m "not gpu"
fi
exec python3 -m pytest -q -p no:cacheprovider "$@"
