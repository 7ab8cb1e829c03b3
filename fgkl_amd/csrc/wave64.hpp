// Wave64 lane moves and wave-local synchronisation for gfx950 (CDNA4, a GFX9 family target, so
// the whole-wave DPP shifts wave_shr:1 / wave_shl:1 are available).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fgkl {

// DPP wave_shr:1 (dpp_ctrl 0x138): lane i receives lane i-1's value; lane 0 keeps `old`.
__device__ __forceinline__ int shr1i(int old, int src) {
  return __builtin_amdgcn_update_dpp(old, src, 0x138, 0xF, 0xF, false);
}
// DPP wave_shl:1 (dpp_ctrl 0x130): lane i receives lane i+1's value; lane 63 keeps `old`.
__device__ __forceinline__ int shl1i(int old, int src) {
  return __builtin_amdgcn_update_dpp(old, src, 0x130, 0xF, 0xF, false);
}
__device__ __forceinline__ uint32_t shr1(uint32_t old, uint32_t src) {
  return uint32_t(shr1i(int(old), int(src)));
}
__device__ __forceinline__ float shr1(float old, float src) {
  return __builtin_bit_cast(float, shr1i(__builtin_bit_cast(int, old), __builtin_bit_cast(int, src)));
}
__device__ __forceinline__ double shr1(double old, double src) {
  const uint64_t o = __builtin_bit_cast(uint64_t, old), s = __builtin_bit_cast(uint64_t, src);
  const uint32_t lo = uint32_t(shr1i(int(uint32_t(o)), int(uint32_t(s))));
  const uint32_t hi = uint32_t(shr1i(int(uint32_t(o >> 32)), int(uint32_t(s >> 32))));
  return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}
__device__ __forceinline__ float shl1(float old, float src) {
  return __builtin_bit_cast(float, shl1i(__builtin_bit_cast(int, old), __builtin_bit_cast(int, src)));
}
__device__ __forceinline__ double shl1(double old, double src) {
  const uint64_t o = __builtin_bit_cast(uint64_t, old), s = __builtin_bit_cast(uint64_t, src);
  const uint32_t lo = uint32_t(shl1i(int(uint32_t(o)), int(uint32_t(s))));
  const uint32_t hi = uint32_t(shl1i(int(uint32_t(o >> 32)), int(uint32_t(s >> 32))));
  return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}

// DPP wave_shr:1 with lane 0 receiving zero (bound_ctrl): no "old" operand to materialise.
__device__ __forceinline__ float shr1z(float src) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, src), 0x138,
                                                            0xF, 0xF, true));
}
__device__ __forceinline__ double shr1z(double src) {
  const uint64_t s = __builtin_bit_cast(uint64_t, src);
  const uint32_t lo = uint32_t(__builtin_amdgcn_mov_dpp(int(uint32_t(s)), 0x138, 0xF, 0xF, true));
  const uint32_t hi =
      uint32_t(__builtin_amdgcn_mov_dpp(int(uint32_t(s >> 32)), 0x138, 0xF, 0xF, true));
  return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}

// DPP wave_shl:1 with lane 63 receiving zero (bound_ctrl).
__device__ __forceinline__ uint32_t shl1z(uint32_t src) {
  return uint32_t(__builtin_amdgcn_mov_dpp(int(src), 0x130, 0xF, 0xF, true));
}
__device__ __forceinline__ float shl1z(float src) {
  return __builtin_bit_cast(float, int(shl1z(__builtin_bit_cast(uint32_t, src))));
}
__device__ __forceinline__ double shl1z(double src) {
  const uint64_t s = __builtin_bit_cast(uint64_t, src);
  const uint64_t lo = shl1z(uint32_t(s)), hi = shl1z(uint32_t(s >> 32));
  return __builtin_bit_cast(double, (hi << 32) | lo);
}

__device__ __forceinline__ float readlane(float v, int l) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ double readlane(double v, int l) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(u)), l));
  const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(u >> 32)), l));
  return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}

// Orders a wave's LDS accesses before and after this point across its lanes.
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Makes a wave's stores visible to the loads it issues next, in any lane: waits until its vector
// memory and LDS operations, stores included, have completed (s_waitcnt vmcnt(0) lgkmcnt(0)), so
// a record written through an LDS pointer and read back through a generic one is there too.  The
// wave's CU shares one vector L1, so no cache maintenance is needed; a workgroup-scope fence
// alone emits no wait on gfx950, and the fences keep the compiler from moving memory operations
// across.
__device__ __forceinline__ void wave_sync_global() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0x0070);  // vmcnt(0) lgkmcnt(0); expcnt unconstrained (gfx9 form)
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

}  // namespace fgkl
