#!/bin/sh
# Builds the C++ examples against the in-tree library (tests/cpp/_build/flow_example, stage_batches).
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
mkdir -p "$HERE/_build"
for prog in flow_example stage_batches; do
    g++ -std=c++20 -O2 -Wall -D__HIP_PLATFORM_AMD__ -I"$ROOT/include" -I/opt/rocm/include "$HERE/$prog.cpp" \
        -L"$ROOT/graphflow_amd" -lgraphflow_amd -L/opt/rocm/lib -lamdhip64 \
        -Wl,-rpath,"\$ORIGIN/../../../graphflow_amd" -Wl,-rpath,/opt/rocm/lib -o "$HERE/_build/$prog"
done
