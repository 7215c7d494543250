// Formats not yet built report LESPMV_ERR_UNSUPPORTED.
#include "common.hpp"

namespace lsp {
[[noreturn]] static void nyi(const char* what) { throw Error(LESPMV_ERR_UNSUPPORTED, std::string(what) + " is not available"); }
void convert_coo(Matrix&, const lespmv_params_t&, hipStream_t) { nyi("COO"); }
void coo_spmv(Matrix&, int, double, const void*, double, void*, hipStream_t) { nyi("COO"); }
void coo_spmm(Matrix&, int, int, double, const void*, int64_t, double, void*, int64_t, hipStream_t) { nyi("COO"); }
void convert_sliced(Matrix&, int, const lespmv_params_t&, hipStream_t) { nyi("sliced ELL"); }
void sliced_spmv(Matrix&, int, double, const void*, double, void*, hipStream_t) { nyi("sliced ELL"); }
void sliced_spmm(Matrix&, int, int, double, const void*, int64_t, double, void*, int64_t, hipStream_t) { nyi("sliced ELL"); }
void sliced_export(Matrix&, int, const std::string&, std::vector<unsigned char>&) { nyi("sliced ELL"); }
void convert_dia(Matrix&, const lespmv_params_t&, hipStream_t) { nyi("DIA"); }
void dia_spmv(Matrix&, int, double, const void*, double, void*, hipStream_t) { nyi("DIA"); }
void dia_spmm(Matrix&, int, int, double, const void*, int64_t, double, void*, int64_t, hipStream_t) { nyi("DIA"); }
void convert_bsr(Matrix&, const lespmv_params_t&, hipStream_t) { nyi("BSR"); }
void bsr_spmv(Matrix&, int, double, const void*, double, void*, hipStream_t) { nyi("BSR"); }
void bsr_spmm(Matrix&, int, int, double, const void*, int64_t, double, void*, int64_t, hipStream_t) { nyi("BSR"); }
void convert_csr5(Matrix&, const lespmv_params_t&, hipStream_t) { nyi("CSR5"); }
void csr5_spmv(Matrix&, double, const void*, double, void*, hipStream_t) { nyi("CSR5"); }
void csr5_export(Matrix&, const std::string&, std::vector<unsigned char>&) { nyi("CSR5"); }
}  // namespace lsp
