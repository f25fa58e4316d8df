#!/bin/sh
# Build the Fortran boundary with amdflang against the in-tree
# libcliffs_gpu.so.  `build.sh test` also builds tests/fortran/fdriver, whose
# own MASTER/PARAMETERS modules stand in for the driver's.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
PKG=$(dirname "$HERE")
ROOT=$(dirname "$PKG")
FC=${FC_FLANG:-/opt/rocm/bin/amdflang}
OUT="$HERE/_build"
mkdir -p "$OUT"
cd "$OUT"
$FC -O2 -fPIC -c "$HERE/cliffs_gpu_mod.f90" -o cliffs_gpu_mod.o
if [ "$1" = "test" ]; then
  $FC -O2 -c "$ROOT/tests/fortran/fdriver_modules.f90" -o fdriver_modules.o
  $FC -O2 -c "$HERE/cliffs_dropin.f90" -o cliffs_dropin.o
  $FC -O2 -c "$ROOT/tests/fortran/fdriver.f90" -o fdriver.o
  $FC -O2 -o "$ROOT/tests/fortran/fdriver" cliffs_gpu_mod.o fdriver_modules.o cliffs_dropin.o \
      fdriver.o -L"$PKG" -lcliffs_gpu -Wl,-rpath,"\$ORIGIN/../../cliffs-src_amd"
  $FC -O2 -c "$ROOT/tests/fortran/fnest.f90" -o fnest.o
  $FC -O2 -o "$ROOT/tests/fortran/fnest" cliffs_gpu_mod.o fdriver_modules.o cliffs_dropin.o \
      fnest.o -L"$PKG" -lcliffs_gpu -Wl,-rpath,"\$ORIGIN/../../cliffs-src_amd"
fi
