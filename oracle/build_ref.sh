#!/usr/bin/env bash
# ORACLE -- test infrastructure only.  Builds the reference solver's own modules, unmodified,
# from where they lie under /root/reference, together with this project's driver
# (ref_harness.f90), into oracle/_ref/.  Only the fixture writer in tests/golden/ runs the
# result; nothing in the product uses it.
#
# Toolchain: amdflang (ROCm) + conda MPICH (mpif.h, libmpifort).  The reference compiles with
# no -O flag (run_compile.sh:5-7), i.e. unoptimised; -ffp-contract=off keeps every multiply and
# add separate as that build does.
set -euo pipefail
REF=${REF:-/root/reference}
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
# REF_OPT: the optimisation level; -O0 (default) is the reference's own build and the one the
# fixtures come from, another level (e.g. -O2, for timing only) goes to oracle/_ref/opt<level>/
REF_OPT=${REF_OPT:--O0}
OUT="$HERE/_ref"
if [ "$REF_OPT" != "-O0" ]; then OUT="$HERE/_ref/opt${REF_OPT#-}"; fi
BLD="$OUT/build"
MPI_PREFIX=${MPI_PREFIX:-/opt/conda}
FC=${FC:-amdflang}

if [ ! -d "$REF" ]; then
  echo "build_ref: $REF not present; skipping" >&2
  exit 0
fi
mkdir -p "$BLD"
FFLAGS="$REF_OPT -ffp-contract=off -I$MPI_PREFIX/include -module-dir $BLD"
# dependency order of the modules PROGRAM ModSolver uses (run_compile.sh:5-7 lists the files)
MODS="ModuleVariables ModPLOT3D_IO ModDecomp ModCommunicate ModInputOutput ModAllocate ModAssign \
ModOperator ModDeriv ModTGfilter ModShockCapturing ModControlForcing ModBC ModPatch ModAxis ModRHS \
ModRungeKutta"
OBJS=""
for m in $MODS; do
  $FC $FFLAGS -c "$REF/$m.f90" -o "$BLD/$m.o"
  OBJS="$OBJS $BLD/$m.o"
done
gcc -O0 -c "$REF/plot3d_format.c" -o "$BLD/plot3d_format.o"
$FC $FFLAGS -c "$HERE/ref_harness.f90" -o "$BLD/ref_harness.o"
$FC -o "$OUT/ref_harness" "$BLD/ref_harness.o" $OBJS "$BLD/plot3d_format.o" \
  -L"$MPI_PREFIX/lib" -Wl,-rpath,"$MPI_PREFIX/lib" -lmpifort -lmpi
echo "build_ref: built $OUT/ref_harness"
