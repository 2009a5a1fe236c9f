#!/bin/bash
# Builds oracle/_ref/libref.so from the reference's own MPI-only translation
# units (params, proc, gridsize, communicator, tensor0, tensor1, particle) as
# they lie under /root/reference, plus oracle/ref_harness.cpp.  grid.cpp,
# poisson.cpp, poisson_IO.cpp and simulation.cpp include the Sandia FFT and
# Hypre headers, which are absent, so they are not built.  Output goes only to
# oracle/_ref/ (git-ignored).  TEST INFRASTRUCTURE.
#
# MPICH is taken from /opt/conda.  No rpath is recorded: the loader of this
# library (tests/oracle_lib.py) opens /opt/conda/lib/libmpi.so.12 by path
# first, so the system libstdc++/libgomp stay the ones in use.
set -euo pipefail
REF=${REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
MPIROOT=${MPIROOT:-/opt/conda}
if [ ! -d "$REF" ] || [ ! -f "$MPIROOT/include/mpi.h" ]; then
  echo "build_ref: reference sources or MPI headers not present; skipping" >&2
  exit 0
fi
mkdir -p "$HERE/_ref"
SRCS="params.cpp proc.cpp gridsize.cpp communicator.cpp tensor0.cpp tensor1.cpp particle.cpp"
ARGS=()
for s in $SRCS; do ARGS+=("$REF/$s"); done
g++ -O3 -ffp-contract=off -fPIC -shared -w -I"$REF" -I"$MPIROOT/include" \
  "${ARGS[@]}" "$HERE/ref_harness.cpp" -L"$MPIROOT/lib" -lmpi -o "$HERE/_ref/libref.so"
echo "build_ref: wrote $HERE/_ref/libref.so"
