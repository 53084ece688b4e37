#!/bin/bash
# Builds the reference solver from its own sources for one flow case
# (test infrastructure; build container only).
#
#   build_ref.sh <build_dir> <n_cubes> <U0> <mu_L>
#
# The reference translation units are compiled where they lie under the
# reference tree; the two configuration headers come from gen_ref_cfg.py in
# <build_dir>/cfg, found first through "-I<cfg> -I-".  The driver
# ref_driver.cpp replaces the reference main program.  Every output goes to
# <build_dir>, which lives under oracle/_ref/.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${BCM_REFERENCE:-/root/reference}"
MPI_ROOT="${BCM_MPI_ROOT:-/opt/conda}"
OUT="$1"
NCUBES="$2"
U0="$3"
MUL="$4"
mkdir -p "$OUT/cfg" "$OUT/obj"
python3 "$HERE/gen_ref_cfg.py" "$REF" "$OUT/cfg" "$NCUBES" "$U0" "$MUL"
# "-I-" also stops quoted includes inside the C++ library headers from
# finding their own directory; the parallel-algorithms headers need theirs.
SEARCH="$(g++ -x c++ -E -v /dev/null 2>&1 || true)"
GXXINC="$(printf '%s\n' "$SEARCH" | grep -E '^ /.*/c\+\+/[0-9]+$' | head -n1 | tr -d ' ' || true)"
CXXFLAGS="-O2 -ffp-contract=off -fno-fast-math -fopenmp -w -I$OUT/cfg -I- -I$REF -I$MPI_ROOT/include -I$GXXINC/pstl"
TUS="BCM_Flux_XYZ_Viscous_Runge_kutta BCM_Ghostcell_minus BCM_Ghostcell_plus
     BCM_X_boundary_condition BCM_Y_boundary_condition BCM_Z_boundary_condition
     BCM_Interface BCM_Interface_table BCM_Grid BCM_FWS_Interface BCM_Initial_condition
     BCM_Immersed_boundary BCM_Reading_IBM BCM_Output"
pids=()
for tu in $TUS; do
  g++ $CXXFLAGS -c "$REF/$tu.cpp" -o "$OUT/obj/$tu.o" 2> "$OUT/obj/$tu.log" &
  pids+=($!)
done
g++ $CXXFLAGS -c "$HERE/ref_driver.cpp" -o "$OUT/obj/ref_driver.o"
for p in "${pids[@]}"; do wait "$p"; done
g++ -fopenmp -o "$OUT/ref_driver" "$OUT"/obj/*.o -L"$MPI_ROOT/lib" -Wl,-rpath,"$MPI_ROOT/lib" -lmpicxx -lmpi
echo "built $OUT/ref_driver"
