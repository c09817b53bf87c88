#!/bin/sh
# Compiles the reference's DM-QR path from its own sources, where they lie
# under /root/reference, into oracle/_ref/libqrdm_ref.so (git-ignored).
# Test infrastructure only: the compiled reference pins the oracle and gives
# the CPU baseline.  Nothing is copied out of /root/reference.
#
# The reference includes its vendored LAPACKE/CBLAS headers and links a
# system BLAS/LAPACK/LAPACKE (setup_QRDM.py:13,38).  The BLAS/LAPACK present
# in this image is OpenBLAS 0.3.28 as bundled with scipy, which exports the
# same functions under a "scipy_" prefix; the -D flags below map the names at
# compile time.  No header, source or library is written for the build.
set -e
REF=${QRDM_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -f "$REF/src/dgeqrdm_work.c" ]; then
  echo "build_ref.sh: reference sources not found under $REF; skipping" >&2
  exit 0
fi
BLASDIR=$(python3 -c "import scipy, os; print(os.path.realpath(os.path.join(os.path.dirname(scipy.__file__), '..', 'scipy.libs')))")
BLAS=$(ls "$BLASDIR"/libscipy_openblas-*.so | head -n 1)
mkdir -p "$OUT"
MAP=""
for f in cblas_dnrm2 cblas_dsyrk cblas_dgemv cblas_dger cblas_dscal cblas_dswap \
         LAPACKE_dlarft LAPACKE_dlarfb_work LAPACKE_dgeqrf LAPACKE_dormqr LAPACKE_dgeqpf \
         LAPACKE_dlamch LAPACKE_dge_nancheck LAPACKE_dtr_nancheck LAPACKE_lsame \
         LAPACKE_xerbla xerbla_ dlapy2_ iladlc_ iladlr_; do
  MAP="$MAP -D$f=scipy_$f"
done
# Sources compiled by the reference's own build list (setup_QRDM.py:20-31),
# minus the Python wrapper (it needs <python3.8/Python.h>) and the unused
# dgeqp3.c wrapper.
gcc -O2 -fPIC -shared -w -std=gnu99 -I"$REF/include" $MAP \
  "$REF/src/dgeqrdm.c" "$REF/src/dgeqrdm_work.c" "$REF/src/dgeqr2.c" \
  "$REF/src/dlarfg.c" "$REF/src/dlarf.c" "$REF/src/dlarfb.c" \
  -o "$OUT/libqrdm_ref.so" "$BLAS" -Wl,-rpath,"$BLASDIR" -lm
echo "$BLAS" > "$OUT/blas_path.txt"
echo "built $OUT/libqrdm_ref.so against $BLAS"
