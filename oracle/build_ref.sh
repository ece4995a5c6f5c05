#!/bin/bash
# Compiles the reference's own CPU simulator (single-threaded STSim path) from
# its sources under /root/reference, together with ref_harness.cpp, into
# oracle/_ref/ref_stsim_f32 (real = float) and ref_stsim_f64 (-DUSE_DOUBLE).
# Test infrastructure only.  Needs nothing beyond g++ and the reference tree;
# exits 0 without building when the reference tree is absent.
set -euo pipefail
REF=${BSIM_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF/src" ]; then
    echo "reference tree not found at $REF; skipping" >&2
    exit 0
fi
mkdir -p "$OUT"
SRC="$REF/src"
UNITS="$SRC/base/NeuronBase.cpp $SRC/base/PopulationBase.cpp $SRC/base/SimInfo.cpp \
       $SRC/base/SimulatorBase.cpp $SRC/sim/SingleThreadSimulator.cpp \
       $SRC/neuron/lif/LIFENeuron.cpp $SRC/neuron/array/ArrayNeuron.cpp \
       $SRC/synapse/static/StaticSynapse.cpp $SRC/net/Network.cpp \
       $SRC/third_party/json/json.cpp $SRC/utils/utils.cpp"
# -include cmath: Network.cpp calls round() without including <cmath>.
FLAGS="-std=c++11 -O3 -w -include cmath -I$SRC"
for prec in f32 f64; do
    extra=""
    [ "$prec" = f64 ] && extra="-DUSE_DOUBLE"
    g++ $FLAGS $extra -o "$OUT/ref_stsim_$prec" "$HERE/ref_harness.cpp" $UNITS
done
echo "built $OUT/ref_stsim_f32 $OUT/ref_stsim_f64"
