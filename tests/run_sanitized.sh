#!/bin/bash
# The CPU test suite (-m "not gpu") against a build of libgeneig.so whose host
# code carries AddressSanitizer and UndefinedBehaviorSanitizer (make
# SANITIZE=1; device code is not instrumented).  Python and the other libraries
# are not instrumented, so the ASan runtime is preloaded and leak checking is
# off; any ASan or UBSan report aborts the run with a non-zero status.
# Usage: tests/run_sanitized.sh [extra pytest arguments]
set -euo pipefail
ROOT=$(cd "$(dirname "$0")/.." && pwd)
make -s -C "$ROOT/sunny-batched-geneig_amd" SANITIZE=1 -j8
ASAN_RT=$(/opt/rocm/llvm/bin/clang -print-file-name=libclang_rt.asan-x86_64.so)
export GENEIG_LIBRARY="$ROOT/sunny-batched-geneig_amd/_build_san/libgeneig.so"
export ASAN_OPTIONS=detect_leaks=0:abort_on_error=1:halt_on_error=1
export UBSAN_OPTIONS=print_stacktrace=1:halt_on_error=1
cd "$ROOT"
LD_PRELOAD="$ASAN_RT" python -m pytest tests -m "not gpu" -x -q -p no:cacheprovider "$@"
