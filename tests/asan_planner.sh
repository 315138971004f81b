#!/bin/bash
# Host planner under AddressSanitizer: builds the FP64 QuEST library with
# -fsanitize=address (g++, host code only) into a temporary directory beside
# a copy of the built engine library, then plans many circuits through it.
# Run from the repository root after building (dgquest_amd/lib).
set -e
D=$(mktemp -d)
H=dgquest_amd/csrc/host
cp dgquest_amd/lib/libdgq_engine.so "$D/"
g++ -O1 -g -std=c++17 -fPIC -fsanitize=address -fno-omit-frame-pointer -Iinclude -I$H -DQuEST_PREC=2 -shared \
    -o "$D/libQuEST_f64.so" $H/quest.cpp $H/dag.cpp $H/planner.cpp $H/mt19937.cpp $H/serve.cpp $H/tune.cpp \
    $H/dense.cpp $H/paging.cpp -L"$D" -ldgq_engine -ldl -Wl,-Bsymbolic -Wl,-rpath,'$ORIGIN'
ASAN_OPTIONS=detect_leaks=0 LD_PRELOAD=$(gcc -print-file-name=libasan.so) python3 tests/asan_plans.py "$D"
rm -rf "$D"
