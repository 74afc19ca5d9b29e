#!/usr/bin/env bash
# Builds rv-sparse's own scalar library from its sources, where they lie, into
# oracle/_ref/librvsp_ref.so. It is a second CPU checker next to oracle.c and
# may serve as the timed CPU baseline. Only the reference's C sources are
# compiled (no reference build system, no generated code, no stand-ins); the
# three RISC-V Vector accumulators are left out exactly as rv-sparse's
# Makefile does for a non-vector -march. Its flags (-std=c11 -O3 -flto),
# with -march=x86-64-v3 instead of -march=native: the library is built here
# and runs on another host, so it targets any AVX2 host.
# Skips quietly (exit 0) when the reference tree is absent.
set -euo pipefail
REF=${RVSP_REFERENCE_DIR:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF/src" ]; then
    echo "build_ref: $REF not present, skipping"
    exit 0
fi
mkdir -p "$OUT/obj"
CFLAGS="-std=c11 -O3 -flto -march=x86-64-v3 -fPIC -I$REF/include -w"
objs=()
while IFS= read -r src; do
    case "$src" in
        */accum_rvv_f32.c|*/accum_contig_f32.c|*/accum_adaptive_f32.c) continue ;;
    esac
    obj="$OUT/obj/$(echo "${src#$REF/}" | tr '/' '_' ).o"
    gcc $CFLAGS -c "$src" -o "$obj"
    objs+=("$obj")
done < <(find "$REF/src" -name '*.c' | sort)
gcc -O3 -flto -march=x86-64-v3 -shared -Wl,-Bsymbolic -o "$OUT/librvsp_ref.so.tmp" "${objs[@]}" -lm
mv "$OUT/librvsp_ref.so.tmp" "$OUT/librvsp_ref.so"
# rv-sparse's Matrix Market reader (tools/src: the formatter and its vector),
# the checker of rvsp_mtx_load
gcc -std=c11 -O2 -fPIC -shared -w -I"$REF/tools/include" \
    "$REF/tools/src/mtx_to_csr_formatter.c" "$REF/tools/src/vec.c" \
    -o "$OUT/librvsp_mtx_ref.so.tmp"
mv "$OUT/librvsp_mtx_ref.so.tmp" "$OUT/librvsp_mtx_ref.so"
# rv-sparse's synthetic generator (tools/src/genmat.c, single-threaded: the
# OpenMP pragmas are ignored without -fopenmp, as in its default build), the
# checker of genmat_generate_csr
gcc -std=c11 -O3 -fPIC -shared -w -I"$REF/tools/include" \
    "$REF/tools/src/genmat.c" -o "$OUT/librvsp_genmat_ref.so.tmp" -lm
mv "$OUT/librvsp_genmat_ref.so.tmp" "$OUT/librvsp_genmat_ref.so"
rm -rf "$OUT/obj"
echo "build_ref: wrote $OUT/librvsp_ref.so, librvsp_mtx_ref.so and librvsp_genmat_ref.so"
# rv-sparse's own C callers of rv_sparse.h -- its two API tests and three
# examples -- compiled unchanged from where they lie, twice each:
#   callers/<name>.hip  against this build's include/ and linked to
#                       rv-sparse_amd/lib/librvsparse_hip.so (the drop-in)
#   callers/<name>.ref  against rv-sparse's include/ and linked to
#                       _ref/librvsp_ref.so (its own scalar library)
# Run side by side, the pair shows that a caller needs no source change.
HIPLIB="$HERE/../rv-sparse_amd/lib"
if [ -f "$HIPLIB/librvsparse_hip.so" ]; then
    mkdir -p "$OUT/callers"
    for src in "$REF/tests/test_spgemm_reference_f32.c" "$REF/tests/test_spgemm_descr_f32.c" \
               "$REF/examples/spgemm_csr_f32.c" "$REF/examples/spgemm_csr_descr_f32.c" \
               "$REF/examples/spgemm_csr_reuse_f32.c"; do
        name=$(basename "$src" .c)
        gcc -std=c11 -O2 -w -I"$HERE/../include" -I"$REF/tests" "$src" \
            -L"$HIPLIB" -lrvsparse_hip -Wl,-rpath,'$ORIGIN/../../../rv-sparse_amd/lib' -lm \
            -o "$OUT/callers/$name.hip"
        gcc -std=c11 -O2 -w -I"$REF/include" -I"$REF/tests" "$src" \
            -L"$OUT" -l:librvsp_ref.so -Wl,-rpath,'$ORIGIN/..' -lm \
            -o "$OUT/callers/$name.ref"
    done
    echo "build_ref: wrote $OUT/callers (rv-sparse's tests and examples, against both libraries)"
fi
