// Reads the feature values the reference's own extractor (class MTX,
// src/sparse_features.cpp) computes for a MatrixMarket file, for the golden
// fixtures. TEST INFRASTRUCTURE ONLY; compiled by build_ref.sh together with
// the reference sources where they lie.
//
// The MTX members are private and the class has no getters for most of them,
// so this one translation unit opens the class's access for reading: every
// header the class declaration depends on is included first, then the class
// header with `private` read as `public`. Nothing else is changed.
#include <cmath>
#include <iostream>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "general_config.h"
#include "sparse_format.h"
#include "timer.h"
#define private public
#include "sparse_features.h"
#undef private

#include "named.hpp"

namespace {

template <class V>
void put_features(NamedArrays& r, MTX<int, V>& x) {
  auto f = [&](const char* name, double v) { r.put(std::string("f:") + name, &v, 1); };
  f("num_rows", x.num_rows); f("num_cols", x.num_cols); f("num_nnzs", x.num_nnzs); f("nnz_mtx", x.nnz_mtx_);
  f("nnz_lower", x.nnz_lower_); f("nnz_upper", x.nnz_upper_); f("nnz_diagonal", x.nnz_diagonal_);
  f("complete_ndiags", x.complete_ndiags);
  f("min_nnz_each_row", x.min_nnz_each_row_); f("max_nnz_each_row", x.max_nnz_each_row_);
  f("min_nnz_each_col", x.min_nnz_each_col_); f("max_nnz_each_col", x.max_nnz_each_col_);
  f("is_symmetric", x.is_symmetric_ ? 1.0 : 0.0);
  f("pattern_symm", x.pattern_symm_); f("value_symm", x.value_symm_);
  f("nnz_ratio", x.nnz_ratio_); f("diag_close_ratio", x.diag_close_ratio_); f("distance_per_row", x.distance_per_row_);
  f("nz_row_ratio", x.nz_row_ratio_); f("ave_nnz_each_row", x.ave_nnz_each_row_);
  f("var_nnz_each_row", x.var_nnz_each_row_); f("standard_dev_row", x.standard_dev_row_);
  f("Gini_row", x.Gini_row_); f("P_ratio_row", x.P_ratio_row_);
  f("nz_col_ratio", x.nz_col_ratio_); f("ave_nnz_each_col", x.ave_nnz_each_col_);
  f("var_nnz_each_col", x.var_nnz_each_col_); f("standard_dev_col", x.standard_dev_col_);
  f("Gini_col", x.Gini_col_); f("P_ratio_col", x.P_ratio_col_);
  f("diagonal_dominant_ratio", x.diagonal_dominant_ratio_);
  f("max_value_offdiag", x.max_value_offdiag_); f("max_value_diagonal", x.max_value_diagonal_);
  f("row_variability", x.row_variability_); f("col_variability", x.col_variability_);
  f("t_num_blocks", x.t_num_blocks); f("t_num_RB", x.t_num_RB); f("t_num_CB", x.t_num_CB);
  f("t_mod_RB", x.t_mod_RB); f("t_mod_CB", x.t_mod_CB);
  f("t_ave_nnz_all_tiles", x.t_ave_nnz_all_tiles); f("t_ave_nnz_RB", x.t_ave_nnz_RB); f("t_ave_nnz_CB", x.t_ave_nnz_CB);
  f("t_var_nnz_all_tiles", x.t_var_nnz_all_tiles); f("t_var_nnz_RB", x.t_var_nnz_RB); f("t_var_nnz_CB", x.t_var_nnz_CB);
  f("t_standard_dev_all_tiles", x.t_standard_dev_all_tiles); f("t_standard_dev_RB", x.t_standard_dev_RB);
  f("t_standard_dev_CB", x.t_standard_dev_CB);
  f("t_min_nnz_all_tiles", x.t_min_nnz_all_tiles_); f("t_max_nnz_all_tiles", x.t_max_nnz_all_tiles_);
  f("t_min_nnz_each_RB", x.t_min_nnz_each_RB_); f("t_max_nnz_each_RB", x.t_max_nnz_each_RB_);
  f("t_min_nnz_each_CB", x.t_min_nnz_each_CB_); f("t_max_nnz_each_CB", x.t_max_nnz_each_CB_);
  f("t_Gini_all_tiles", x.t_Gini_all_tiles_); f("t_Gini_RB", x.t_Gini_RB_); f("t_Gini_CB", x.t_Gini_CB_);
  f("t_P_ratio_all_tiles", x.t_P_ratio_all_tiles_); f("t_P_ratio_RB", x.t_P_ratio_RB_);
  f("t_P_ratio_CB", x.t_P_ratio_CB_);
  f("t_nz_ratio_tiles", x.t_nz_ratio_tiles_); f("t_nz_ratio_RB", x.t_nz_ratio_RB_); f("t_nz_ratio_CB", x.t_nz_ratio_CB_);
  f("uniqR", x.uniqR); f("uniqC", x.uniqC); f("potReuseR", x.potReuseR); f("potReuseC", x.potReuseC);
  f("GrX_uniqR", x.GrX_uniqR); f("GrX_uniqC", x.GrX_uniqC); f("GrX_potReuseR", x.GrX_potReuseR);
  f("GrX_potReuseC", x.GrX_potReuseC);
  r.put("img:tile_nnz", x.nnz_by_Tiles_); r.put("img:tile_max", x.max_rownnz_per_tile_);
  r.put("img:tile_ave", x.ave_rownnz_per_tile_); r.put("img:tile_std", x.std_rownnz_per_tile_);
  r.put("img:tile_uniq_rows", x.uniq_RB); r.put("img:tile_uniq_cols", x.uniq_CB);
  r.put("img:rb_nnz", x.nnz_by_RB_); r.put("img:rb_max", x.max_rownnz_per_RB_);
  r.put("img:rb_ave", x.ave_rownnz_per_RB_); r.put("img:rb_std", x.std_rownnz_per_RB_);
  r.put("img:cb_nnz", x.nnz_by_CB_); r.put("img:cb_max", x.max_colnnz_per_CB_);
  r.put("img:cb_ave", x.ave_colnnz_per_CB_); r.put("img:cb_std", x.std_colnnz_per_CB_);
  r.put("img:row_nnz", x.nnz_by_row_); r.put("img:col_nnz", x.nnz_by_col_);
}

}  // namespace

// Runs Get_features' sequence (test/Get_features.cpp:46-54) on one file with
// IndexType int and ValueType double.
extern "C" int ref_features(void* res, const char* path) {
  auto& r = *static_cast<NamedArrays*>(res);
  MTX<int, double> x(0);
  if (!x.MtxLoad(path)) return 1;
  x.CalculateFeatures();
  x.CalculateTilesExtraFeatures(path);
  put_features(r, x);
  return 0;
}
