// Raw buffer loads and stores for the column passes.
//
// A buffer access takes a uniform 128-bit resource (base address and size, held
// in SGPRs), a per-lane 32-bit byte offset (one VGPR) and a uniform byte offset
// (an SGPR).  A column-pass thread reads its E elements at rows t + q T of one
// column, so the per-lane offset (t * width + c) * bytes stays the same for the
// whole kernel and the row term q T width and the tile's first column move into
// the SGPR offset: one VGPR of addressing per thread instead of a 64-bit address
// per element kept live across the transforms.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ptx {

// Word 3 of a raw (untyped, stride-0) buffer descriptor on the gfx9 family,
// gfx950 included: 32-bit data format, no swizzle.
constexpr int kBufferWord3 = 0x00020000;

// Resource for `bytes` bytes at `base` (bytes < 2^31); accesses whose offset
// reaches `bytes` read zero and drop stores.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, kBufferWord3);
}

__device__ __forceinline__ float2 buf_load_f2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}

__device__ __forceinline__ float buf_load_f1(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}

__device__ __forceinline__ void buf_store_f2(float2 v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  using u2 = __attribute__((ext_vector_type(2))) unsigned;
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, (int)voff, (int)soff, 0);
}

__device__ __forceinline__ void buf_store_f1(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff, 0);
}

}  // namespace ptx
