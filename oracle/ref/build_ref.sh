#!/usr/bin/env bash
# Build recipe for oracle/_ref: the reference Fortran (sbrdar/fesom2_tracer_dwarf, lib/*.F90 and
# src/fesom_analytic.F90) compiled in place from $FESOM_REF with AMD flang in FP64
# (-fdefault-real-8, the reference's "dp" build), plus our own driver oracle/ref/ref_driver.F90.
#
# This is test infrastructure only: the products are a checker (golden-vector generator) and a
# CPU baseline, never part of the shipped path.  Nothing from the reference is copied into the
# repository; all outputs go to oracle/_ref/ (git-ignored).
#
# The reference uses `use mpi`.  The image ships MPICH (libmpi, mpif.h) under /opt/conda, but its
# mpi.mod was written by gfortran and cannot be read by flang, so the `mpi` module is compiled here
# from that same MPICH mpif.h (no stub: every MPI call resolves to the image's libmpi).
#
# Usage: oracle/ref/build_ref.sh            (FESOM_REF defaults to /root/reference)
set -euo pipefail

HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
REPO="$(cd "$HERE/../.." && pwd)"
REF="${FESOM_REF:-/root/reference}"
OUT="$REPO/oracle/_ref"
MPI_ROOT="${MPI_ROOT:-/opt/conda}"
FC="${FC:-flang}"

if [ ! -d "$REF/lib" ]; then
    echo "build_ref: reference sources not found at $REF (set FESOM_REF)" >&2
    exit 3
fi
if ! command -v "$FC" >/dev/null 2>&1; then
    echo "build_ref: Fortran compiler '$FC' not found" >&2
    exit 3
fi
if [ ! -f "$MPI_ROOT/include/mpif.h" ]; then
    echo "build_ref: $MPI_ROOT/include/mpif.h not found" >&2
    exit 3
fi

mkdir -p "$OUT/obj" "$OUT/mod" "$OUT/inc"
cp "$MPI_ROOT/include/mpif.h" "$OUT/inc/"
cat > "$OUT/inc/mpi_module.F90" <<'EOF'
module mpi
  implicit none
  include 'mpif.h'
end module mpi
EOF

# -O2 -fdefault-real-8: the reference dp build.  No -march: the x86-64 baseline has no FMA, so the
# compiler cannot contract a*b+c; -ffp-contract=off states it explicitly.
FFLAGS=(-cpp -O2 -fdefault-real-8 -ffp-contract=off -I"$OUT/inc" -module-dir "$OUT/mod" -I"$OUT/mod")

"$FC" -c "${FFLAGS[@]}" "$OUT/inc/mpi_module.F90" -o "$OUT/obj/mpi_module.o"

ORDER="hp_math_intrinsics oce_modules gen_modules_config gen_modules_rotate_grid fortran_utils
MOD_READ_BINARY_ARRAYS MOD_WRITE_BINARY_ARRAYS fesom_profiler MOD_MESH MOD_PARTIT MOD_TRACER MOD_DYN
io_restart_derivedtype mesh_output gen_modules_partitioning gen_halo_exchange oce_mesh
oce_muscl_adv oce_adv_tra_fct oce_adv_tra_hor oce_adv_tra_ver oce_adv_tra_driver analytic_mesh
tracer_init_from_mesh tracer_c_interface"

OBJS=("$OUT/obj/mpi_module.o")
for f in $ORDER; do
    src="$REF/lib/$f.F90"
    obj="$OUT/obj/$f.o"
    if [ ! -f "$obj" ] || [ "$src" -nt "$obj" ]; then
        "$FC" -c "${FFLAGS[@]}" "$src" -o "$obj"
    fi
    OBJS+=("$obj")
done

LIBS=(-L"$MPI_ROOT/lib" -Wl,-rpath,"$MPI_ROOT/lib" -lmpifort -lmpi)

# The reference's own analytic driver (reproduces its published console golden).
"$FC" -c "${FFLAGS[@]}" "$REF/src/fesom_analytic.F90" -o "$OUT/obj/fesom_analytic.o"
"$FC" "$OUT/obj/fesom_analytic.o" "${OBJS[@]}" "${LIBS[@]}" -o "$OUT/fesom_tracer_analytic"

# Our driver: reads a mesh/field bundle, runs the reference kernels, writes golden vectors.
"$FC" -c "${FFLAGS[@]}" "$HERE/ref_driver.F90" -o "$OUT/obj/ref_driver.o"
"$FC" "$OUT/obj/ref_driver.o" "${OBJS[@]}" "${LIBS[@]}" -o "$OUT/ref_driver"

# OpenMP variant of the same driver for the CPU timing baseline (the reference's own OpenMP
# regions; its lock-based scatter is not bitwise reproducible, so it is never used as a checker).
OMPFLAGS=(-cpp -O2 -fdefault-real-8 -ffp-contract=off -fopenmp -I"$OUT/inc" -module-dir "$OUT/mod_omp"
          -I"$OUT/mod_omp")
mkdir -p "$OUT/obj_omp" "$OUT/mod_omp"
"$FC" -c "${OMPFLAGS[@]}" "$OUT/inc/mpi_module.F90" -o "$OUT/obj_omp/mpi_module.o"
OMP_OBJS=("$OUT/obj_omp/mpi_module.o")
for f in $ORDER; do
    src="$REF/lib/$f.F90"
    obj="$OUT/obj_omp/$f.o"
    if [ ! -f "$obj" ] || [ "$src" -nt "$obj" ]; then
        "$FC" -c "${OMPFLAGS[@]}" "$src" -o "$obj"
    fi
    OMP_OBJS+=("$obj")
done
"$FC" -c "${OMPFLAGS[@]}" "$HERE/ref_driver.F90" -o "$OUT/obj_omp/ref_driver.o"
OMP_LIB_DIR="${OMP_LIB_DIR:-/opt/rocm/llvm/lib}"
"$FC" -fopenmp "$OUT/obj_omp/ref_driver.o" "${OMP_OBJS[@]}" "${LIBS[@]}" -Wl,-rpath,"$OMP_LIB_DIR" \
    -o "$OUT/ref_driver_omp"

echo "build_ref: built $OUT/fesom_tracer_analytic, $OUT/ref_driver and $OUT/ref_driver_omp"
