#!/bin/bash
# Builds the parts of the reference that compile from their own sources into
# oracle/_ref/ (test infrastructure; never shipped or linked by the product):
#   libref_rcm.so  source/rcm.cpp (Burkardt GENRCM, std headers only) + ref_rcm_shim.cpp
# Everything else on the reference's path is CUDA or needs HOOMD headers and
# is not buildable here.
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${FSD_REFERENCE:-/root/reference}"
[ -f "$REF/source/rcm.cpp" ] || { echo "reference sources not found under $REF; nothing built"; exit 0; }
mkdir -p "$HERE/_ref"
g++ -O2 -fPIC -shared -w -I"$REF/source" -o "$HERE/_ref/libref_rcm.so" "$REF/source/rcm.cpp" "$HERE/ref_rcm_shim.cpp"
echo "built $HERE/_ref/libref_rcm.so"
