// MatrixMarket input and COO -> CSR.
//
// lespmv_read_mtx restates read_coo_matrix / read_csr_matrix
// (utils/sparse_io.cpp:26-204): coordinate real/integer/pattern matrices,
// 1-based -> 0-based, pattern entries read as 1.0, values read as double and
// converted, symmetric matrices expanded in file order (each off-diagonal
// entry followed by its mirror), then coo_to_csr (include/sparse_conversion.h:
// 29-91): a stable counting sort by row, and with `compact` the duplicate sum of
// sum_csr_duplicates (include/sparse_operation.h:21-73), which drops zero sums
// and lists a row's columns in reverse order of first appearance.
// Errors come back as status codes instead of exit().
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"

namespace lsp {

extern thread_local std::string g_last_error;

namespace {

struct CooHost {
  int64_t m = 0, n = 0;
  bool symmetric = false;
  std::vector<int32_t> ri, ci;
  std::vector<double> v;
};

std::string lower(std::string s) {
  for (auto& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

CooHost read_coo(const char* path) {
  FILE* f = std::fopen(path, "r");
  LSP_REQUIRE(f, LESPMV_ERR_ARG, std::string("Unable to open file: ") + path);
  std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
  char line[4096];
  LSP_REQUIRE(std::fgets(line, sizeof line, f), LESPMV_ERR_ARG, "Could not process Matrix Market banner.");
  std::istringstream banner{std::string(line)};
  std::string tag, object, format, field, symmetry;
  banner >> tag >> object >> format >> field >> symmetry;
  LSP_REQUIRE(tag == "%%MatrixMarket", LESPMV_ERR_ARG, "Could not process Matrix Market banner.");
  object = lower(object);
  format = lower(format);
  field = lower(field);
  symmetry = lower(symmetry);
  LSP_REQUIRE(object == "matrix", LESPMV_ERR_ARG, "Invalid Matrix");
  LSP_REQUIRE(format == "coordinate" && (field == "real" || field == "integer" || field == "pattern"),
              LESPMV_ERR_UNSUPPORTED, "Only sparse real-valued or pattern coordinate matrices are supported");
  // skip comments, then the size line
  do {
    LSP_REQUIRE(std::fgets(line, sizeof line, f), LESPMV_ERR_ARG, "The line of rows, cols, nnzs is in wrong format");
  } while (line[0] == '%');
  long long rows, cols, nnz;
  LSP_REQUIRE(std::sscanf(line, "%lld %lld %lld", &rows, &cols, &nnz) == 3, LESPMV_ERR_ARG,
              "The line of rows, cols, nnzs is in wrong format");
  CooHost c;
  c.m = rows;
  c.n = cols;
  c.ri.resize(nnz);
  c.ci.resize(nnz);
  c.v.resize(nnz);
  const bool pattern = field == "pattern";
  for (long long i = 0; i < nnz; i++) {
    long long r, cc;
    double val = 1.0;
    const int got = pattern ? std::fscanf(f, "%lld %lld", &r, &cc) : std::fscanf(f, "%lld %lld %lf", &r, &cc, &val);
    LSP_REQUIRE(got == (pattern ? 2 : 3), LESPMV_ERR_ARG, "truncated MatrixMarket entry list");
    LSP_REQUIRE(r >= 1 && r <= rows && cc >= 1 && cc <= cols, LESPMV_ERR_ARG, "MatrixMarket index out of range");
    c.ri[i] = (int32_t)(r - 1);
    c.ci[i] = (int32_t)(cc - 1);
    c.v[i] = val;
  }
  if (symmetry == "symmetric") {
    c.symmetric = true;
    std::vector<int32_t> ri, ci;
    std::vector<double> v;
    ri.reserve(2 * nnz);
    ci.reserve(2 * nnz);
    v.reserve(2 * nnz);
    for (long long i = 0; i < nnz; i++) {
      ri.push_back(c.ri[i]);
      ci.push_back(c.ci[i]);
      v.push_back(c.v[i]);
      if (c.ri[i] != c.ci[i]) {
        ri.push_back(c.ci[i]);
        ci.push_back(c.ri[i]);
        v.push_back(c.v[i]);
      }
    }
    c.ri.swap(ri);
    c.ci.swap(ci);
    c.v.swap(v);
  }
  LSP_REQUIRE((int64_t)c.v.size() < INT32_MAX, LESPMV_ERR_UNSUPPORTED, "sizes must fit 32-bit indices");
  return c;
}

// Host coo_to_csr with optional compact (the reference algorithm, serial).
void coo_to_csr_host(const CooHost& c, bool compact, std::vector<int32_t>& rp, std::vector<int32_t>& ci,
                     std::vector<double>& v) {
  const int64_t m = c.m, nnz = (int64_t)c.v.size();
  rp.assign(m + 1, 0);
  ci.resize(nnz);
  v.resize(nnz);
  for (int64_t i = 0; i < nnz; i++) rp[c.ri[i]]++;
  int64_t acc = 0;
  for (int64_t i = 0; i < m; i++) {
    const int64_t t = rp[i];
    rp[i] = (int32_t)acc;
    acc += t;
  }
  rp[m] = (int32_t)nnz;
  std::vector<int32_t> pos(rp.begin(), rp.end() - 1);
  for (int64_t i = 0; i < nnz; i++) {
    const int32_t d = pos[c.ri[i]]++;
    ci[d] = c.ci[i];
    v[d] = c.v[i];
  }
  if (!compact) return;
  std::vector<int32_t> next(c.n, -1);
  std::vector<double> sums(c.n, 0.0);
  int32_t out = 0, row_end = 0;
  for (int64_t i = 0; i < m; i++) {
    int32_t head = -2;
    const int32_t row_start = row_end;
    row_end = rp[i + 1];
    for (int32_t jj = row_start; jj < row_end; jj++) {
      const int32_t j = ci[jj];
      sums[j] += v[jj];
      if (next[j] == -1) {
        next[j] = head;
        head = j;
      }
    }
    while (head != -2) {
      const int32_t cur = head;
      head = next[cur];
      if (sums[cur] != 0) {
        ci[out] = cur;
        v[out] = sums[cur];
        out++;
      }
      next[cur] = -1;
      sums[cur] = 0;
    }
    rp[i + 1] = out;
  }
  ci.resize(out);
  v.resize(out);
}

void fill_out(lespmv_host_csr_t* out, int dtype, int64_t m, int64_t n, const std::vector<int32_t>& rp,
              const std::vector<int32_t>& ci, const std::vector<double>& v) {
  out->m = m;
  out->n = n;
  out->nnz = (int64_t)ci.size();
  out->dtype = dtype;
  out->mm_flags = 0;
  out->row_offset = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * (m + 1)));
  out->col_index = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * std::max<size_t>(ci.size(), 1)));
  out->values = std::malloc(vsize(dtype) * std::max<size_t>(v.size(), 1));
  LSP_REQUIRE(out->row_offset && out->col_index && out->values, LESPMV_ERR_OOM, "host allocation failed");
  std::memcpy(out->row_offset, rp.data(), sizeof(int32_t) * (m + 1));
  if (!ci.empty()) std::memcpy(out->col_index, ci.data(), sizeof(int32_t) * ci.size());
  for (size_t k = 0; k < v.size(); k++) {
    if (dtype == LESPMV_F64) static_cast<double*>(out->values)[k] = v[k];
    else static_cast<float*>(out->values)[k] = (float)v[k];
  }
}

// ---- GPU coo_to_csr without compaction: stable radix sort of row keys ----

__global__ void iota_kernel(int64_t n, uint32_t* __restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (uint32_t)i;
}

template <class V>
__global__ void gather_entries(int64_t n, const uint32_t* __restrict__ perm, const int32_t* __restrict__ ci_in,
                               const V* __restrict__ v_in, int32_t* __restrict__ ci_out, V* __restrict__ v_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ci_out[i] = ci_in[perm[i]];
  v_out[i] = v_in[perm[i]];
}

__global__ void row_counts(int64_t n, const uint32_t* __restrict__ rows, int64_t* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(reinterpret_cast<unsigned long long*>(cnt + rows[i]), 1ull);
}

int bits_for(uint64_t v) {
  int b = 0;
  while (b < 64 && (v >> b)) b++;
  return b;
}

}  // namespace

// Device COO (row, col, val) -> CSR arrays in new device buffers (stable in
// input order inside a row, like the reference's counting sort).
void coo_to_csr_device(int dtype, int64_t m, int64_t nnz, const int32_t* ri, const int32_t* ci, const void* v,
                       DBuf& rp_out, DBuf& ci_out, DBuf& v_out, hipStream_t s) {
  DBuf keys, perm, cnt;
  keys.alloc(sizeof(uint32_t) * std::max<int64_t>(nnz, 1));
  perm.alloc(sizeof(uint32_t) * std::max<int64_t>(nnz, 1));
  cnt.alloc(sizeof(int64_t) * (m + 1));
  LSP_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int64_t) * (m + 1), s));
  if (nnz) {
    LSP_HIP(hipMemcpyAsync(keys.p, ri, sizeof(uint32_t) * nnz, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(iota_kernel, grid1d(nnz, 256), dim3(256), 0, s, nnz, perm.as<uint32_t>());
    LSP_LAUNCHED();
    radix_sort_pairs(keys.as<uint32_t>(), perm.as<uint32_t>(), nnz, bits_for((uint64_t)std::max<int64_t>(m - 1, 0)),
                     s);
    hipLaunchKernelGGL(row_counts, grid1d(nnz, 256), dim3(256), 0, s, nnz, keys.as<uint32_t>(), cnt.as<int64_t>());
    LSP_LAUNCHED();
  }
  const int64_t total = scan_exclusive_i64(cnt.as<int64_t>(), m, s);
  LSP_HIP(hipMemcpyAsync(cnt.as<int64_t>() + m, &total, sizeof(int64_t), hipMemcpyHostToDevice, s));
  auto h = to_host<int64_t>(cnt.p, m + 1, s);
  std::vector<int32_t> h32(h.begin(), h.end());
  rp_out.alloc(sizeof(int32_t) * (m + 1));
  LSP_HIP(hipMemcpyAsync(rp_out.p, h32.data(), sizeof(int32_t) * (m + 1), hipMemcpyHostToDevice, s));
  ci_out.alloc(sizeof(int32_t) * nnz);
  v_out.alloc(vsize(dtype) * nnz);
  if (nnz) {
    if (dtype == LESPMV_F64)
      hipLaunchKernelGGL((gather_entries<double>), grid1d(nnz, 256), dim3(256), 0, s, nnz, perm.as<uint32_t>(), ci,
                         static_cast<const double*>(v), ci_out.as<int32_t>(), v_out.as<double>());
    else
      hipLaunchKernelGGL((gather_entries<float>), grid1d(nnz, 256), dim3(256), 0, s, nnz, perm.as<uint32_t>(), ci,
                         static_cast<const float*>(v), ci_out.as<int32_t>(), v_out.as<float>());
    LSP_LAUNCHED();
  }
  LSP_HIP(hipStreamSynchronize(s));
}

}  // namespace lsp

using namespace lsp;

extern "C" int lespmv_read_mtx(const char* path, int dtype, int compact, lespmv_host_csr_t* out) {
  try {
    LSP_REQUIRE(path && out, LESPMV_ERR_ARG, "null argument");
    LSP_REQUIRE(dtype == LESPMV_F32 || dtype == LESPMV_F64, LESPMV_ERR_ARG, "bad dtype");
    CooHost c = read_coo(path);
    std::vector<int32_t> rp, ci;
    std::vector<double> v;
    coo_to_csr_host(c, compact != 0, rp, ci, v);
    fill_out(out, dtype, c.m, c.n, rp, ci, v);
    out->mm_flags = c.symmetric ? LESPMV_MM_SYMMETRIC : 0;
    return LESPMV_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return LESPMV_ERR_ARG;
  }
}

extern "C" int lespmv_coo_to_csr(int dtype, int64_t m, int64_t n, int64_t nnz, const int32_t* row_index,
                                 const int32_t* col_index, const void* values, int compact, lespmv_host_csr_t* out) {
  try {
    LSP_REQUIRE(out && (nnz == 0 || (row_index && col_index && values)), LESPMV_ERR_ARG, "null argument");
    LSP_REQUIRE(dtype == LESPMV_F32 || dtype == LESPMV_F64, LESPMV_ERR_ARG, "bad dtype");
    LSP_REQUIRE(m >= 0 && n >= 0 && nnz >= 0 && m < INT32_MAX && n < INT32_MAX && nnz < INT32_MAX,
                LESPMV_ERR_UNSUPPORTED, "sizes must fit 32-bit indices");
    for (int64_t i = 0; i < nnz; i++)
      LSP_REQUIRE(row_index[i] >= 0 && row_index[i] < m && col_index[i] >= 0 && col_index[i] < n, LESPMV_ERR_ARG,
                  "COO index out of range");
    if (compact) {
      CooHost c;
      c.m = m;
      c.n = n;
      c.ri.assign(row_index, row_index + nnz);
      c.ci.assign(col_index, col_index + nnz);
      c.v.resize(nnz);
      for (int64_t i = 0; i < nnz; i++)
        c.v[i] = dtype == LESPMV_F64 ? static_cast<const double*>(values)[i] : static_cast<const float*>(values)[i];
      std::vector<int32_t> rp, ci;
      std::vector<double> v;
      coo_to_csr_host(c, true, rp, ci, v);
      fill_out(out, dtype, m, n, rp, ci, v);
      return LESPMV_OK;
    }
    // GPU path: stable radix sort by row on the device
    hipStream_t s = nullptr;
    DBuf dri, dci, dv, rp, ci, v;
    dri.alloc(sizeof(int32_t) * nnz);
    dci.alloc(sizeof(int32_t) * nnz);
    dv.alloc(vsize(dtype) * nnz);
    if (nnz) {
      LSP_HIP(hipMemcpy(dri.p, row_index, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
      LSP_HIP(hipMemcpy(dci.p, col_index, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
      LSP_HIP(hipMemcpy(dv.p, values, vsize(dtype) * nnz, hipMemcpyHostToDevice));
    }
    coo_to_csr_device(dtype, m, nnz, dri.as<int32_t>(), dci.as<int32_t>(), dv.p, rp, ci, v, s);
    out->m = m;
    out->n = n;
    out->nnz = nnz;
    out->dtype = dtype;
    out->mm_flags = 0;
    out->row_offset = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * (m + 1)));
    out->col_index = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * std::max<int64_t>(nnz, 1)));
    out->values = std::malloc(vsize(dtype) * std::max<int64_t>(nnz, 1));
    LSP_REQUIRE(out->row_offset && out->col_index && out->values, LESPMV_ERR_OOM, "host allocation failed");
    LSP_HIP(hipMemcpy(out->row_offset, rp.p, sizeof(int32_t) * (m + 1), hipMemcpyDeviceToHost));
    if (nnz) {
      LSP_HIP(hipMemcpy(out->col_index, ci.p, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost));
      LSP_HIP(hipMemcpy(out->values, v.p, vsize(dtype) * nnz, hipMemcpyDeviceToHost));
    }
    return LESPMV_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return LESPMV_ERR_ARG;
  }
}
