// benchmark_spmv: the reference's per-format drivers (test/benchmark_spmv_<fmt>.cpp)
// in one executable on the MI355X library. Reads a MatrixMarket file with
// read_csr_matrix, runs test_<fmt>_matrix_kernels (conversion, comparison
// against CSR, device-timed benchmark) and prints one result line per format:
//   <matID> <matrix> <FORMAT> <kernel_tag> <ms> <GFLOP/s> <convert_ms>
// Options: --format=csr|coo|ell|sell|sell_c_sigma|sell_c_R|dia|bsr|csr5|all
//          --precision=32|64  --kernel=<tag>  --matID=<id>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/lespmv/LeSpMV.h"

static std::string arg(int argc, char** argv, const char* key, const char* def) {
  const size_t n = std::strlen(key);
  for (int i = 1; i < argc; i++)
    if (std::strncmp(argv[i], key, n) == 0 && argv[i][n] == '=') return argv[i] + n + 1;
  return def;
}

template <class V>
static int run(const char* path, const std::string& fmt, int kernel, int mat_id) {
  auto csr = read_csr_matrix<int, V>(path);
  std::string name(path);
  name = name.substr(name.find_last_of('/') + 1);
  printf("matrix %s: %d x %d, %d nonzeros, %s\n", name.c_str(), csr.num_rows, csr.num_cols, csr.num_nnzs,
         sizeof(V) == 8 ? "double" : "float");
  std::vector<std::string> fmts;
  if (fmt == "all") fmts = {"csr", "coo", "ell", "sell", "sell_c_sigma", "sell_c_R", "dia", "bsr", "csr5"};
  else fmts = {fmt};
  for (const auto& f : fmts) {
    double ms = 0, conv = 0;
    try {
      if (f == "csr") ms = test_csr_matrix_kernels(csr, kernel, 0);
      else if (f == "coo") ms = test_coo_matrix_kernels(csr, kernel, 0, conv);
      else if (f == "ell") ms = test_ell_matrix_kernels(csr, kernel, RowMajor, 0, conv);
      else if (f == "sell") ms = test_s_ell_matrix_kernels(csr, kernel, 0, conv);
      else if (f == "sell_c_sigma") ms = test_sell_c_sigma_matrix_kernels(csr, kernel, 0, conv);
      else if (f == "sell_c_R") ms = test_sell_c_R_matrix_kernels(csr, kernel, 0);
      else if (f == "dia") ms = test_dia_matrix_kernels(csr, kernel, 0, conv);
      else if (f == "bsr") ms = test_bsr_matrix_kernels(csr, kernel, 0, conv);
      else if (f == "csr5") ms = test_csr5_matrix_kernels<int, uint32_t, V>(csr, kernel, 0);
      else {
        fprintf(stderr, "unknown format %s\n", f.c_str());
        return 2;
      }
    } catch (const lespmv::Error& e) {
      printf("%d %s %s %d skipped (%s)\n", mat_id, name.c_str(), f.c_str(), kernel, e.what());
      continue;
    }
    const double gflops = ms > 0 ? 2.0 * csr.num_nnzs / (ms * 1e-3) / 1e9 : 0.0;
    printf("%d %s %s %d %8.4f %5.4f %8.4f\n", mat_id, name.c_str(), f.c_str(), kernel, ms, gflops, conv);
  }
  delete_host_matrix(csr);
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2 || argv[1][0] == '-') {
    fprintf(stderr, "usage: %s <matrix.mtx> [--format=...] [--precision=32|64] [--kernel=K] [--matID=N]\n", argv[0]);
    return 1;
  }
  const std::string fmt = arg(argc, argv, "--format", "all");
  const int prec = std::stoi(arg(argc, argv, "--precision", "64"));
  const int kernel = std::stoi(arg(argc, argv, "--kernel", "0"));
  const int mat_id = std::stoi(arg(argc, argv, "--matID", "0"));
  try {
    lespmv::check(lespmv_init(0), "lespmv_init");
    return prec == 32 ? run<float>(argv[1], fmt, kernel, mat_id) : run<double>(argv[1], fmt, kernel, mat_id);
  } catch (const std::exception& e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 3;
  }
}
