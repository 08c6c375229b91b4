#!/bin/sh
# Builds the gfx950 device library (libnicp_amd.so) and the host-only
# synthetic-scene library (libnicp_synth.so) in place.
set -e
here="$(cd "$(dirname "$0")" && pwd)"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
"$HIPCC" --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared -ffp-contract=off \
    -Xclang -target-feature -Xclang -packed-fp32-ops \
    -Wall -Wno-unused-function -Wno-unused-variable \
    -o "$here/libnicp_amd.so.tmp" "$here/csrc/nicp_capi.hip"
mv "$here/libnicp_amd.so.tmp" "$here/libnicp_amd.so"
g++ -O3 -fopenmp -std=gnu++17 -fPIC -shared -Wall \
    -o "$here/libnicp_synth.so.tmp" "$here/csrc/synth.cpp"
mv "$here/libnicp_synth.so.tmp" "$here/libnicp_synth.so"
g++ -O2 -std=c++17 -Wall -I "$here/../include" -o "$here/nicp_simple_aligner.tmp" "$here/apps/nicp_simple_aligner.cpp" \
    -L "$here" -lnicp_amd -Wl,-rpath,'$ORIGIN'
mv "$here/nicp_simple_aligner.tmp" "$here/nicp_simple_aligner"
g++ -O2 -std=c++17 -Wall -I "$here/../include" -o "$here/nicp_aligner.tmp" "$here/apps/nicp_aligner.cpp" \
    -L "$here" -lnicp_amd -Wl,-rpath,'$ORIGIN'
mv "$here/nicp_aligner.tmp" "$here/nicp_aligner"
