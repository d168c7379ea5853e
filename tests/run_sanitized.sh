#!/bin/bash
# The CPU test suite with the host code of libprr.so and the oracle built with
# AddressSanitizer + UndefinedBehaviorSanitizer (clang's runtime; device code
# is not instrumented): the .params parser, the PNG / JPEG / PPM decoders,
# the host-run integrator and the C ABI's argument checks all run under it.
# Builds into parrelray_amd/lib-asan and oracle/_asan; extra arguments go to pytest.
set -euo pipefail
ROOT=$(cd "$(dirname "$0")/.." && pwd)
make -C "$ROOT/parrelray_amd/csrc" -j8 OUT=../lib-asan SANITIZE=1 ../lib-asan/libprr.so
make -C "$ROOT/oracle" asan
RT="$(/opt/rocm/llvm/bin/clang++ -print-resource-dir)/lib/linux/libclang_rt.asan-x86_64.so"
export PRR_LIB="$ROOT/parrelray_amd/lib-asan/libprr.so"
export PRR_ORACLE_LIB="$ROOT/oracle/_asan/liboracle.so"
export ASAN_OPTIONS=detect_leaks=0:halt_on_error=1:abort_on_error=1
export UBSAN_OPTIONS=halt_on_error=1:print_stacktrace=1
cd "$ROOT"
LD_PRELOAD="$RT" python -m pytest tests -q -m "not gpu" -p no:cacheprovider "$@"
