#!/usr/bin/env bash
# Compile the reference LeSpMV sources, where they lie under /root/reference,
# together with oracle/ref_driver.cpp into oracle/_ref/libref_lespmv.so.
# Checker only: used in this container to write tests/golden fixtures and to
# cross-check the oracle. Nothing is copied out of the reference.
#
# The reference's CMakeLists forces icx; its CSR5 kernel uses AVX-512
# intrinsics that g++ 11 rejects, so ROCm's clang++ (with libomp) is used.
set -euo pipefail
REF=${REF:-/root/reference/LeSpMV}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF/src" ]; then
  echo "reference sources not present at $REF; skipping" >&2
  exit 0
fi
CXX=${REF_CXX:-/opt/rocm/lib/llvm/bin/clang++}
mkdir -p "$OUT"
SRCS=(
  "$REF/src/spmv_csr.cpp" "$REF/src/spmv_coo.cpp" "$REF/src/spmv_ell.cpp" "$REF/src/spmv_s_ell.cpp"
  "$REF/src/spmv_sell_c_sigma.cpp" "$REF/src/spmv_sell_c_R.cpp" "$REF/src/spmv_dia.cpp" "$REF/src/spmv_bsr.cpp"
  "$REF/src/spmv_csr5.cpp" "$REF/utils/thread.cpp" "$REF/utils/partition.cpp"
  "$REF/utils/sparse_io.cpp" "$REF/utils/mmio.cpp" "$REF/src/sparse_features.cpp"
)
"$CXX" -O2 -std=c++17 -fPIC -shared -fopenmp -march=native -w \
  -I"$REF/include" -I"$HERE" \
  "$HERE/ref_driver.cpp" "$HERE/ref_features.cpp" "${SRCS[@]}" \
  -Wl,-rpath,/opt/rocm/lib/llvm/lib -o "$OUT/libref_lespmv.so"
echo "built $OUT/libref_lespmv.so"
# the CPU baseline (ref_bench.cpp over the reference's CSR kernel), in two
# portable instruction levels instead of -march=native
for LEVEL in x86-64-v3 x86-64-v4; do
  "$CXX" -O3 -std=c++17 -fPIC -shared -fopenmp -march=$LEVEL -w \
    -I"$REF/include" "$HERE/ref_bench.cpp" "$REF/src/spmv_csr.cpp" "$REF/utils/thread.cpp" "$REF/utils/partition.cpp" \
    -Wl,-rpath,/opt/rocm/lib/llvm/lib -o "$OUT/libref_csr_bench_$LEVEL.so"
  echo "built $OUT/libref_csr_bench_$LEVEL.so"
done
