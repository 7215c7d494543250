// get_features: the reference's Get_features (test/Get_features.cpp) on the
// MI355X library. Reads a MatrixMarket file, computes every MTX feature on the
// GPU, prints them in FeaturesPrint / ExtraFeaturesPrint's layout, and prints
//   <matID> <matrix> <feature ms>
// (the line Get_features appends to its features log). With --write=<root> it
// also writes <root>/<name>/<name>.features and the tile/block images.
// Options: --matID=<id>  --precision=32|64  --write=<root>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../../include/lespmv/LeSpMV.h"

static std::string arg(int argc, char** argv, const char* key, const char* def) {
  const size_t n = std::strlen(key);
  for (int i = 1; i < argc; i++)
    if (std::strncmp(argv[i], key, n) == 0 && argv[i][n] == '=') return argv[i] + n + 1;
  return def;
}

template <class V>
static int run(const char* path, int mat_id, const std::string& root) {
  MTX<int, V> mtx(mat_id);
  mtx.MtxLoad(path);
  mtx.CalculateFeatures();
  mtx.CalculateTilesExtraFeatures(path);
  mtx.FeaturesPrint();
  mtx.ExtraFeaturesPrint();
  printf("%d %s %5.4f\n", mat_id, mtx.getMatName().c_str(), mtx.features().device_ms);
  if (!root.empty() && !mtx.FeaturesWrite(root.c_str())) {
    fprintf(stderr, "error: could not write features under %s\n", root.c_str());
    return 3;
  }
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2 || argv[1][0] == '-') {
    fprintf(stderr, "usage: %s my_matrix.mtx [--matID=N] [--precision=32|64] [--write=<root>]\n", argv[0]);
    return 1;
  }
  try {
    lespmv::check(lespmv_init(0), "lespmv_init");
    const int mat_id = std::stoi(arg(argc, argv, "--matID", "0"));
    const int prec = std::stoi(arg(argc, argv, "--precision", "64"));
    const std::string root = arg(argc, argv, "--write", "");
    return prec == 32 ? run<float>(argv[1], mat_id, root) : run<double>(argv[1], mat_id, root);
  } catch (const std::exception& e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 3;
  }
}
