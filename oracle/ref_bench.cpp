// CPU baseline from the reference's own code: LeSpMV_csr with kernel_flag = 2
// (src/spmv_csr.cpp:164-185, the load-balanced OpenMP kernel the reference
// driver uses, test/benchmark_spmv_csr.cpp:109), its partition computed once
// up front as utils/test_spmv_csr.cpp:68-72 does. Compiled by build_ref.sh
// from the sources under the reference tree into oracle/_ref/ (x86-64-v3 and
// x86-64-v4 builds, so the baseline runs on hosts with or without AVX-512).
// Test infrastructure: only bench.py's cpu_baseline leg loads it.
#include <omp.h>

#include <chrono>

#include "sparse_format.h"
#include "sparse_partition.h"
#include "spmv_csr.h"
#include "thread.h"

extern "C" {

// Milliseconds per y = A*x (alpha 1, beta 0) over `iters` calls after one
// untimed call; y receives the product.
double refbench_csr_lb(int m, int n, int nnz, const int* rp, const int* ci, const double* v, const double* x,
                       double* y, int threads, int iters) {
  Le_set_thread_num(threads);
  omp_set_num_threads(threads);
  CSR_Matrix<int, double> A;
  A.num_rows = m;
  A.num_cols = n;
  A.num_nnzs = nnz;
  A.row_offset = const_cast<int*>(rp);
  A.col_index = const_cast<int*>(ci);
  A.values = const_cast<double*>(v);
  A.kernel_flag = 2;
  int* part = new int[threads + 1];
  balanced_partition_row_by_nnz(rp, m, threads, part);
  A.partition = part;
  LeSpMV_csr<int, double>(1.0, A, x, 0.0, y);
  const auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < iters; i++) LeSpMV_csr<int, double>(1.0, A, x, 0.0, y);
  const auto t1 = std::chrono::steady_clock::now();
  delete[] part;
  return std::chrono::duration<double, std::milli>(t1 - t0).count() / (iters > 0 ? iters : 1);
}

}  // extern "C"
