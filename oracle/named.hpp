// Named-array result container shared by the CPU oracle and the reference
// driver. TEST INFRASTRUCTURE ONLY: nothing in the product links this file.
//
// A conversion produces a bag of named, typed host arrays ("reorder",
// "chunk_len", "col_index", ...) plus scalar fields stored as one-element
// int64 arrays. ctypes callers query the byte size of a name and copy it out.
#pragma once
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

struct NamedArrays {
  std::map<std::string, std::vector<unsigned char>> arrays;
  std::string error;

  template <class T>
  void put(const std::string& name, const T* src, size_t n) {
    auto& v = arrays[name];
    v.resize(n * sizeof(T));
    if (n) std::memcpy(v.data(), src, n * sizeof(T));
  }
  template <class T>
  void put(const std::string& name, const std::vector<T>& src) {
    put(name, src.data(), src.size());
  }
  void scalar(const std::string& name, int64_t value) { put(name, &value, 1); }
  template <class T>
  const T* get(const std::string& name) const {
    auto it = arrays.find(name);
    return it == arrays.end() ? nullptr : reinterpret_cast<const T*>(it->second.data());
  }
  int64_t get_scalar(const std::string& name) const {
    const int64_t* p = get<int64_t>(name);
    return p ? *p : -1;
  }
};

#define NAMED_ARRAYS_CABI(prefix)                                              \
  extern "C" void* prefix##_result_new() { return new NamedArrays(); }         \
  extern "C" void prefix##_result_free(void* h) {                              \
    delete static_cast<NamedArrays*>(h);                                       \
  }                                                                            \
  extern "C" int64_t prefix##_result_size(void* h, const char* name) {         \
    auto* r = static_cast<NamedArrays*>(h);                                    \
    auto it = r->arrays.find(name);                                            \
    return it == r->arrays.end() ? -1 : (int64_t)it->second.size();            \
  }                                                                            \
  extern "C" int prefix##_result_get(void* h, const char* name, void* dst) {   \
    auto* r = static_cast<NamedArrays*>(h);                                    \
    auto it = r->arrays.find(name);                                            \
    if (it == r->arrays.end()) return -1;                                      \
    if (!it->second.empty()) std::memcpy(dst, it->second.data(), it->second.size()); \
    return 0;                                                                  \
  }                                                                            \
  extern "C" int prefix##_result_names(void* h, char* dst, int64_t cap) {      \
    auto* r = static_cast<NamedArrays*>(h);                                    \
    std::string all;                                                           \
    for (auto& kv : r->arrays) { all += kv.first; all += ','; }                \
    if ((int64_t)all.size() + 1 > cap) return -1;                              \
    std::memcpy(dst, all.c_str(), all.size() + 1);                             \
    return 0;                                                                  \
  }                                                                            \
  extern "C" const char* prefix##_result_error(void* h) {                      \
    return static_cast<NamedArrays*>(h)->error.c_str();                        \
  }
