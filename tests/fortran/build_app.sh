#!/bin/bash
# Compiles the reference's Weibel, shock and reconnection apps
# (proj/<app>/app.f90 and main.f90, unchanged, read where they lie) against
# this package's drop-in Fortran modules -- boundary_shock and
# boundary_reconnection included, in place of the apps' own -- into
# tests/fortran/_app/<app>_app.x (git-ignored).  Each app then runs its own
# loop on the device through libwmpic_f.so and libwmpic.so.
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
for app in weibel shock reconnection; do
  mkdir -p "$TMP/$app"
  amdflang -O2 -cpp -I "$PKG/fortran/mod" -module-dir "$TMP/$app" \
    "$REF/proj/$app/app.f90" "$REF/proj/$app/main.f90" \
    -o "$OUT/${app}_app.x.tmp" -L "$PKG" -lwmpic_f -lwmpic -Wl,-rpath,'$ORIGIN/../../../wumingpic2d_amd'
  mv "$OUT/${app}_app.x.tmp" "$OUT/${app}_app.x"
  echo "build_app: built $OUT/${app}_app.x"
done
