#!/bin/bash
# Compile check of the Fortran drop-in modules legendre.hip.F90 and
# transform.hip.F90 (with lsd_c.F90 and lsd_dropin.F90) against the
# reference's own parameters / dyn_mpi / meshs / variables modules, and of
# the reference's timestep, velocity, magnetic, codensity, composition,
# rotation and nonlinear modules, unchanged, against the drop-ins (serial
# build, _Nth = _Nr = 1), all compiled where they lie under /root/reference.
# Nothing of the reference is copied: its templates are reached through
# symlinks in the build folder, and all outputs (.mod, .o) stay in
# leeds-code_amd/_build/dropin.  Needs the reference sources and amdflang,
# so it runs in the build container only.  (oracle/ref_build.py --dropin
# links the same set into a running program.)
set -euo pipefail
REF=${REF:-/root/reference}
FC=${FC_DROPIN:-/opt/rocm/bin/amdflang}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/../_build/dropin
rm -rf "$OUT"
mkdir -p "$OUT/inc/sub"
ln -s "$REF/Scripts/parallel.h_template" "$OUT/inc/parallel.h"
ln -s "$REF/Scripts/parameters.F90_template" "$OUT/parameters.F90"
cd "$OUT"
FFLAGS="-cpp -c -O0 -I$OUT/inc/sub -I$OUT/inc"
$FC $FFLAGS parameters.F90
$FC $FFLAGS "$REF/modules/dyn_mpi.F90"
$FC $FFLAGS "$REF/modules/meshs.F90"
$FC $FFLAGS "$REF/modules/variables.F90"
$FC $FFLAGS "$HERE/lsd_c.F90"
$FC $FFLAGS "$HERE/lsd_dropin.F90"
$FC $FFLAGS "$HERE/legendre.hip.F90"
$FC $FFLAGS "$HERE/transform.hip.F90"
for m in timestep velocity magnetic codensity composition rotation nonlinear; do
  $FC $FFLAGS "$REF/modules/$m.F90"
done
echo "legendre.hip.F90 and transform.hip.F90 compile against the reference's parameters and variables modules,"
echo "and the reference's timestep, velocity, magnetic, codensity, composition, rotation and nonlinear modules"
echo "compile unchanged against them"
