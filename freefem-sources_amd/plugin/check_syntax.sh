#!/bin/sh
# Compiles plugin/ffhip.cpp against FreeFEM's own headers (syntax and type
# check only: the FreeFEM executable that would dlopen the plugin is not built
# here).  -DCMAKE is FreeFEM's switch that skips the autoconf config.h.
set -e
REF=${FF_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
[ -d "$REF/src/fflib" ] || { echo "FreeFEM tree not found at $REF" >&2; exit 3; }
g++ -std=gnu++14 -fsyntax-only -DCMAKE -w \
  -I"$REF/src/fflib" -I"$REF/src/femlib" -I"$REF/src/Graphics" -I"$REF/src/bamglib" \
  -I"$REF/src/lglib" -I"$REF/src/Algo" -I"$HERE/../../include" "$HERE/ffhip.cpp"
echo "plugin/ffhip.cpp: OK against $REF headers"
