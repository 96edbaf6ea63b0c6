#!/bin/bash
# Compiles the reference's own particle, sort and moment modules, exactly as they lie in
# the reference tree, into oracle/_ref/libwmref.so.  Nothing else is compiled:
# field.f90 and boundary_periodic.f90 `use mpi`, and this image has no MPI
# module in the compiler's format, so those two files are not buildable here.
# libwmref.so is single-threaded (the bit-exact pins); libwmref_omp.so is the
# same sources with the reference's OpenMP loops on, which bench.py times beside
# the C restatement's same phases.  Both are test infrastructure only.
set -euo pipefail
REF=${WM_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -f "$REF/common/particle.f90" ]; then
  echo "build_ref: no reference tree at $REF; skipping" >&2
  exit 0
fi
mkdir -p "$OUT"
amdflang -O2 -fPIC -shared -module-dir "$OUT" \
  "$REF/common/particle.f90" "$REF/common/sort.f90" "$REF/common/mom_calc.f90" \
  -o "$OUT/libwmref.so.tmp"
mv "$OUT/libwmref.so.tmp" "$OUT/libwmref.so"
mkdir -p "$OUT/omp"
amdflang -O2 -fopenmp -fPIC -shared -module-dir "$OUT/omp" \
  "$REF/common/particle.f90" "$REF/common/sort.f90" "$REF/common/mom_calc.f90" \
  -Wl,-rpath,/opt/rocm/llvm/lib -o "$OUT/libwmref_omp.so.tmp"
mv "$OUT/libwmref_omp.so.tmp" "$OUT/libwmref_omp.so"
echo "build_ref: built $OUT/libwmref.so and $OUT/libwmref_omp.so"
