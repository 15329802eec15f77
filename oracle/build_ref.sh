#!/bin/bash
# Builds oracle/_ref/libref.so: the reference's own sources (2D, SIMD blocking with SIMD_LEN 16, no MPI, PETSc
# or hwloc), compiled unmodified where they lie under REF, plus oracle/ref_harness.cpp.  LAPACK/BLAS come
# from MKL (libmkl_rt) in this image.  Test infrastructure only: nothing in the product loads the result.
# Usage: oracle/build_ref.sh [REF]   (default /root/reference); exits 0 without building when REF or MKL is
# absent.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
REF="${1:-/root/reference}"
MKL_DIR="${MKL_DIR:-/opt/conda/lib}"
OUT="$HERE/_ref"
if [ ! -d "$REF/src" ] || [ ! -e "$MKL_DIR/libmkl_rt.so" ]; then
    echo "build_ref: reference sources or MKL not present; nothing built"
    exit 0
fi
DEFS="-fPIC -fopenmp -fopenmp-simd -DSPACE_DIMS=2 -DENABLE_SIMD_BLOCKING -DSIMD_LEN=16 -DLOGLEVEL=1 -DUSE_ALIGNED_ALLOC -I$REF/src"
compile() {
    local f="$1" o
    o="$OBJ/$(echo "${f#"$REF"/src/}" | tr '/' '_').o"
    if [ "$o" -nt "$f" ]; then return 0; fi
    case "$f" in
        *.c) gcc -std=gnu11 $FLAGS -c "$f" -o "$o" ;;
        *) g++ -std=gnu++14 $FLAGS -c "$f" -o "$o" ;;
    esac
}
export -f compile
# libref.so: -O2 without FMA contraction across x86-64 baseline code, the build the fixtures and parity
# tests compare against; libref_fast.so: the reference's own optimisation flags (-O3, OpenMP SIMD) for
# x86-64-v3 (AVX2 + FMA, which every current EPYC and Xeon host has), timed as the CPU baseline.
for variant in "libref.so:obj:-O2" "libref_fast.so:obj_fast:-O3 -march=x86-64-v3"; do
    LIBNAME="${variant%%:*}"; rest="${variant#*:}"; OBJDIR="${rest%%:*}"; OPT="${rest#*:}"
    OBJ="$OUT/$OBJDIR"
    FLAGS="$OPT $DEFS"
    mkdir -p "$OBJ"
    export OBJ REF FLAGS
    # every source below src/ except the drivers (main_*.cpp, which need generated headers)
    find "$REF/src" -mindepth 2 \( -name '*.cpp' -o -name '*.c' \) | sort | xargs -P "${JOBS:-8}" -I{} bash -c 'compile "$@"' _ {}
    g++ -std=gnu++14 $FLAGS -c "$HERE/ref_harness.cpp" -o "$OBJ/ref_harness.o"
    g++ -shared -fopenmp -o "$OUT/$LIBNAME" "$OBJ"/*.o -L"$MKL_DIR" -lmkl_rt -Wl,-rpath,"$MKL_DIR"
    echo "build_ref: $OUT/$LIBNAME"
done
