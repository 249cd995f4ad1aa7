#!/bin/bash
# Runs the CPU tests of the C ABI (input validation, the host search
# controls, stage one's quasi-Newton, the Wald test) against the host-side
# ASan + UBSan build of the library: edge-rs_amd/build.py --sanitize first.
set -e
ROOT="$(cd "$(dirname "$0")/.." && pwd)"
LIB="$ROOT/edge-rs_amd/_build_asan/libnebula_hip.so"
[ -f "$LIB" ] || python "$ROOT/edge-rs_amd/build.py" --sanitize
RT="$(ls /opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so | head -1)"
export NH_LIBRARY="$LIB"
export ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=1"
export UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1"
LD_PRELOAD="$RT" python -m pytest -q -m "not gpu" -p no:cacheprovider \
  "$ROOT/tests/test_capi_host.py" "$ROOT/tests/test_search_host.py" \
  "$ROOT/tests/test_stage_one_host.py" "$ROOT/tests/test_sc_test.py" "$@"
