#!/bin/sh
# Build the Fortran side with amdflang against the in-tree libcliffs_gpu.so:
#   - the NetCDF-3 shim (../netcdf3: libcliffs_nc3.a for Fortran programs,
#     libcliffs_nc3.so for the Python tests),
#   - the ISO_C_BINDING module and the GPU driver cliffs_gpu_driver (the
#     reference's command line, parameter file and NetCDF files),
#   - with `build.sh test`: tests/fortran/fdriver and fnest, whose own
#     MASTER/PARAMETERS modules stand in for the driver's, and the shim's
#     Fortran test program tests/fortran/fnc3.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
PKG=$(dirname "$HERE")
ROOT=$(dirname "$PKG")
NC3="$PKG/netcdf3"
FC=${FC_FLANG:-/opt/rocm/bin/amdflang}
CC=${CC_HOST:-gcc}
OUT="$HERE/_build"
mkdir -p "$OUT"
cd "$OUT"
$CC -O2 -fPIC -Wall -c "$NC3/nc3.c" -o nc3.o
ar rcs "$NC3/libcliffs_nc3.a" nc3.o
$CC -shared -o "$NC3/libcliffs_nc3.so" nc3.o
$FC -O2 -fPIC -c "$NC3/netcdf3_mod.f90" -o netcdf3_mod.o
$FC -O2 -fPIC -c "$HERE/cliffs_gpu_mod.f90" -o cliffs_gpu_mod.o
# no contraction: host arithmetic feeding the solver rounds as real*8 code
$FC -O2 -ffp-contract=off -c "$HERE/cliffs_driver.f90" -o cliffs_driver.o
$FC -O2 -o "$HERE/cliffs_gpu_driver" cliffs_driver.o cliffs_gpu_mod.o netcdf3_mod.o "$NC3/libcliffs_nc3.a" \
    -L"$PKG" -lcliffs_gpu -Wl,-rpath,"\$ORIGIN/.."
if [ "$1" = "test" ]; then
  $FC -O2 -c "$ROOT/tests/fortran/fdriver_modules.f90" -o fdriver_modules.o
  $FC -O2 -c "$HERE/cliffs_dropin.f90" -o cliffs_dropin.o
  $FC -O2 -c "$ROOT/tests/fortran/fdriver.f90" -o fdriver.o
  $FC -O2 -o "$ROOT/tests/fortran/fdriver" cliffs_gpu_mod.o fdriver_modules.o cliffs_dropin.o \
      fdriver.o -L"$PKG" -lcliffs_gpu -Wl,-rpath,"\$ORIGIN/../../cliffs-src_amd"
  $FC -O2 -c "$ROOT/tests/fortran/fnest.f90" -o fnest.o
  $FC -O2 -o "$ROOT/tests/fortran/fnest" cliffs_gpu_mod.o fdriver_modules.o cliffs_dropin.o \
      fnest.o -L"$PKG" -lcliffs_gpu -Wl,-rpath,"\$ORIGIN/../../cliffs-src_amd"
  $FC -O2 -c "$ROOT/tests/fortran/fnc3.f90" -o fnc3.o
  $FC -O2 -o "$ROOT/tests/fortran/fnc3" fnc3.o netcdf3_mod.o "$NC3/libcliffs_nc3.a"
fi
