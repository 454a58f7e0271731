// Raw buffer loads/stores with hardware range checking (gfx950).
//
// A buffer resource carries the base address and the byte size of one
// array; an access at an offset beyond the size returns zero (loads) or is
// dropped (stores).  Kernels pass kOOB as the offset of lanes or nodes that
// must not touch memory, so memory instructions are never skipped and the
// compiler's wait counts stay exact.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace adt {

using rsrc_t = __amdgpu_buffer_rsrc_t;

constexpr uint32_t kOOB = 0xFFFFFFF0u;  // out-of-range buffer offset

__device__ __forceinline__ rsrc_t make_rsrc(const void* p, uint64_t bytes) {
  const uint32_t nb = bytes >= 0xFFFFFFF0ull ? 0xFFFFFFF0u : (uint32_t)bytes;
  return __builtin_amdgcn_make_buffer_rsrc((void*)p, 0, nb, 0x00020000);
}

__device__ __forceinline__ double bload(rsrc_t r, uint32_t off, double) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ float bload(rsrc_t r, uint32_t off, float) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ void bstore(rsrc_t r, uint32_t off, double v) {
  using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, v), r, off, 0, 0);
}
__device__ __forceinline__ void bstore(rsrc_t r, uint32_t off, float v) {
  using V = decltype(__builtin_amdgcn_raw_buffer_load_b32(r, 0, 0, 0));
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(V, v), r, off, 0, 0);
}

// Byte offset of element `node`, or kOOB when !ok.
template <typename T>
__device__ __forceinline__ uint32_t boff(bool ok, int64_t node) {
  return ok ? (uint32_t)(node * (int64_t)sizeof(T)) : kOOB;
}

}  // namespace adt
