#!/bin/sh
# Builds oracle/_ref/libffref_ref.so: FreeFEM's own femlib sources
# (HashMatrix, CG, QuadratureFormular, and the Mesh3 / FESpace3 / P1-P2
# element sources), libMesh's libmeshb7.c (used by Mesh3dn.cpp) and the
# no-graphics host globals of src/Graphics/gggg.cpp, compiled where they lie
# under the reference tree, plus the test drivers ref_driver.cpp and
# seam_driver.cpp (the plugin's glue, freefem-sources_amd/plugin/
# ffhip_glue.hpp, run against these objects; it links libffhip.so, which must
# be built first).
# -DCMAKE is the reference's own switch that skips the autoconf config.h
# (src/Graphics/rgraph.hpp:55-57).  Flags follow the reference build:
# g++ -O3 -mavx2 -std=gnu++14 (no FMA contraction).
# Test infrastructure only; needs /root/reference (absent on GPU hosts).
set -e
REF=${FF_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
[ -d "$REF/src/femlib" ] || { echo "reference tree not found at $REF" >&2; exit 3; }
mkdir -p "$OUT"
CXXFLAGS="-O3 -mavx2 -std=gnu++14 -fPIC -DCMAKE -DNCHECKPTR -w"
INC="-I$REF/src/femlib -I$REF/src/Graphics -I$REF/src/fflib -I$REF/src/libMesh"
FEMLIB="HashMatrix CG QuadratureFormular Mesh1dn Mesh2dn Mesh3dn MeshSn MeshLn GQuadTree FQuadTree \
  FESpacen fem P012_1d P012_2d P012_3d P012_3dSurf P012_3dCurve mshptg splitsimplex"
for f in $FEMLIB; do
  g++ $CXXFLAGS $INC -c "$REF/src/femlib/$f.cpp" -o "$OUT/$f.o" &
done
gcc -O2 -fPIC -w -I"$REF/src/libMesh" -c "$REF/src/libMesh/libmeshb7.c" -o "$OUT/libmeshb7.o" &
wait
g++ $CXXFLAGS $INC -c "$REF/src/Graphics/gggg.cpp" -o "$OUT/gggg.o"
g++ $CXXFLAGS $INC -c "$HERE/ref_driver.cpp" -o "$OUT/ref_driver.o"
PKG="$HERE/../freefem-sources_amd"
[ -f "$PKG/libffhip.so" ] || { echo "build libffhip.so first (make -C freefem-sources_amd)" >&2; exit 4; }
g++ $CXXFLAGS $INC -I"$HERE/../include" -I"$PKG/plugin" -c "$HERE/seam_driver.cpp" -o "$OUT/seam_driver.o"
OBJS=""
for f in $FEMLIB; do OBJS="$OBJS $OUT/$f.o"; done
g++ -shared -o "$OUT/libffref_ref.so" "$OUT/ref_driver.o" "$OUT/seam_driver.o" $OBJS "$OUT/gggg.o" \
  "$OUT/libmeshb7.o" -L"$PKG" -lffhip -Wl,-rpath,'$ORIGIN/../../freefem-sources_amd'

echo "built $OUT/libffref_ref.so"
