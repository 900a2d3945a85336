#!/bin/sh
# Builds oracle/_ref/libref_path.so from the reference's own C++ sources where
# they lie under /root/reference (nothing is copied into this repository):
# the minimal set that links processSecretion and the .mat snapshot writer
# (output_writer.cpp, matlab_writer.cpp), updateMotility (motility.cpp) and
# tryFormAllAttachments (cell_interactions.cpp)
# with no undefined symbols.
# Test infrastructure only.  Exits 0 without building when the reference is
# absent (the GPU box has no /root/reference).
set -e
REF=${REFERENCE_ROOT:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -f "$REF/src/cell_phenotype.cpp" ]; then
  echo "build_ref: reference sources not present; skipping"
  exit 0
fi
mkdir -p "$OUT"
g++ -std=c++17 -O2 -fPIC -shared -Wl,--no-undefined -w \
  -I"$REF/shaders" -I"$REF/lib/pugixml" -I"$REF/src" \
  "$HERE/ref_driver.cpp" \
  "$REF/src/cell_data.cpp" "$REF/src/cell_phenotype.cpp" \
  "$REF/src/cell_definitions.cpp" "$REF/src/config_parser.cpp" \
  "$REF/src/output_writer.cpp" "$REF/src/matlab_writer.cpp" \
  "$REF/src/motility.cpp" "$REF/src/cell_interactions.cpp" \
  "$REF/lib/pugixml/pugixml.cpp" \
  -o "$OUT/libref_path.so.tmp"
mv "$OUT/libref_path.so.tmp" "$OUT/libref_path.so"
echo "build_ref: $OUT/libref_path.so"
