#!/usr/bin/env bash
# Build the Fortran host (axisymmetricjet_amd/ajet_solver) with amdflang against libajet.so,
# and the PLOT3D round-trip checker used by the CPU tests.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
PKG="$(dirname "$HERE")"
BLD="$PKG/_build/host"
FC=${FC:-amdflang}
mkdir -p "$BLD"
FFLAGS="-O2 -module-dir $BLD"
for f in ajet_capi ajet_plot3d ajet_solver; do
  $FC $FFLAGS -c "$HERE/$f.f90" -o "$BLD/$f.o"
done
$FC -o "$PKG/ajet_solver" "$BLD/ajet_capi.o" "$BLD/ajet_plot3d.o" "$BLD/ajet_solver.o" \
  -L"$PKG" -lajet -Wl,-rpath,'$ORIGIN'
$FC $FFLAGS -c "$HERE/p3d_roundtrip.f90" -o "$BLD/p3d_roundtrip.o"
$FC -o "$BLD/p3d_roundtrip" "$BLD/ajet_plot3d.o" "$BLD/p3d_roundtrip.o"
echo "build_host: built $PKG/ajet_solver"
