#!/bin/sh
# Builds oracle/_ref/ref_solver_driver and oracle/_ref/ref_der_driver: the
# reference's FP64 solver and compact derivatives compiled from its own
# sources where they lie (test and baseline infrastructure; the outputs stay
# in oracle/_ref/, which git ignores).  These modules need only LAPACK/BLAS
# (dgemm, dgesv, dgetrf, dgetri, dgeev), taken from the MKL that the image
# carries under /opt/conda/lib (gfortran LP64 interface, GNU OpenMP threading).
#
#   sh oracle/build_ref.sh            (REF, FC and MKL override the defaults)
set -e
REF=${REF:-/root/reference}
FC=${FC:-/opt/rocm/llvm/bin/amdflang}
MKL=${MKL:-/opt/conda/lib}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
mkdir -p "$OUT"
cd "$OUT"
"$FC" -O3 -c "$REF/src/precision/precision.f90"
"$FC" -O3 -c "$REF/src/class_solver/class_solver_coefficient_3d.f90"
"$FC" -O3 -c "$REF/src/class_solver/class_solver_3d.f90"
"$FC" -O3 -c "$HERE/ref_solver_driver.f90"
"$FC" -O3 -o ref_solver_driver precision.o class_solver_coefficient_3d.o class_solver_3d.o \
  ref_solver_driver.o -L"$MKL" -Wl,-rpath,"$MKL" -lmkl_gf_lp64 -lmkl_gnu_thread -lmkl_core \
  -lgomp -lpthread -lm -ldl
"$FC" -O3 -c "$REF/src/class_derivatives/class_derivatives_coefficient.f90"
"$FC" -O3 -c "$REF/src/class_derivatives/class_derivatives.f90"
"$FC" -O3 -c "$HERE/ref_der_driver.f90"
"$FC" -O3 -o ref_der_driver precision.o class_derivatives_coefficient.o class_derivatives.o \
  ref_der_driver.o -L"$MKL" -Wl,-rpath,"$MKL" -lmkl_gf_lp64 -lmkl_gnu_thread -lmkl_core \
  -lgomp -lpthread -lm -ldl
echo "built $OUT/ref_solver_driver $OUT/ref_der_driver"
