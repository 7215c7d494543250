// Loads through buffer resources (device code only).
//
// A buffer load addresses memory as a scalar base (the resource) plus a
// 32-bit per-lane byte offset plus an optional scalar offset. Against a
// 64-bit pointer per lane this saves the 64-bit address operation of every
// gather and, for strided streams, lets the stride live in a scalar register.
// A resource covers fewer than 2^32 bytes; offsets past its size read 0, so
// callers use these loads only for arrays below that size and keep their own
// bounds checks for correctness.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lsp {

// Word 3 of a raw (untyped, unswizzled) buffer resource on gfx9-family parts.
constexpr int kRawBufferWord3 = 0x00020000;

constexpr int64_t kMaxBufferBytes = 0xFFFFFFFFll;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, int64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)(uint32_t)bytes, kRawBufferWord3);
}

// Element at byte offset voff + soff (soff uniform over the wave).
template <class T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff = 0) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit elements");
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}

// Element i of the array the resource covers.
template <class T>
__device__ __forceinline__ T buf_at(__amdgpu_buffer_rsrc_t r, int64_t i) {
  return buf_load<T>(r, (uint32_t)i * (uint32_t)sizeof(T));
}

}  // namespace lsp
