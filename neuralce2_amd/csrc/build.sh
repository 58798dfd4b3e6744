#!/bin/bash
# Build libneuralce_hip.so for gfx950 (in-tree, next to the package).
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
out="${here}/../libneuralce_hip.so"
# -sink-insts-to-avoid-spills: MachineLICM would hoist the FP64 exp's constants out of the NeuralCE step loop and the
# register allocator then spills them (reloaded one after another in every step's Metropolis test)
/opt/rocm/bin/hipcc --offload-arch=gfx950 -O3 -std=c++20 -fPIC -shared \
    -Xclang -target-feature -Xclang -packed-fp32-ops \
    -mllvm -sink-insts-to-avoid-spills \
    -Wall -Wno-unused-function \
    -o "${out}.tmp" "${here}/neuralce_hip.hip"
mv "${out}.tmp" "${out}"
echo "built ${out}"
