// Formats not yet built report LESPMV_ERR_UNSUPPORTED.
#include "common.hpp"

namespace lsp {
[[noreturn]] static void nyi(const char* what) { throw Error(LESPMV_ERR_UNSUPPORTED, std::string(what) + " is not available"); }
void convert_csr5(Matrix&, const lespmv_params_t&, hipStream_t) { nyi("CSR5"); }
void csr5_spmv(Matrix&, double, const void*, double, void*, hipStream_t) { nyi("CSR5"); }
void csr5_export(Matrix&, const std::string&, std::vector<unsigned char>&) { nyi("CSR5"); }
}  // namespace lsp
