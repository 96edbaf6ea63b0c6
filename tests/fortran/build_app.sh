#!/bin/bash
# Compiles the reference's Weibel app (proj/weibel/app.f90 and main.f90,
# unchanged, read where they lie) against this package's drop-in Fortran
# modules, into tests/fortran/_app/weibel_app.x (git-ignored).  The app then
# runs its own loop on the device through libwmpic_f.so and libwmpic.so.
set -euo pipefail
REF=${WM_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
PKG="$HERE/../../wumingpic2d_amd"
if [ ! -f "$REF/proj/weibel/app.f90" ]; then
  echo "build_app: no reference tree at $REF; skipping" >&2
  exit 0
fi
OUT="$HERE/_app"
mkdir -p "$OUT"
TMP=$(mktemp -d)
trap 'rm -rf "$TMP"' EXIT
amdflang -O2 -cpp -I "$PKG/fortran/mod" -module-dir "$TMP" \
  "$REF/proj/weibel/app.f90" "$REF/proj/weibel/main.f90" \
  -o "$OUT/weibel_app.x.tmp" -L "$PKG" -lwmpic_f -lwmpic -Wl,-rpath,'$ORIGIN/../../../wumingpic2d_amd'
mv "$OUT/weibel_app.x.tmp" "$OUT/weibel_app.x"
echo "build_app: built $OUT/weibel_app.x"
