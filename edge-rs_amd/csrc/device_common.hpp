// Wavefront-level helpers for gfx950 (64-lane waves).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nh {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Butterfly sum over the 64 lanes of a wave; every lane gets the total.
// The pairing is fixed, so the result is the same on every run.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
  return v;
}

__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __ballot(p); }

// Values every lane of the wave holds alike, moved to scalar registers
// (arguments of out-of-line device functions arrive in vector registers, so
// the compiler cannot tell they are uniform); uniform base pointers let
// loads use a scalar base with 32-bit per-lane offsets.
__device__ __forceinline__ uint32_t uni(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uni(uint64_t v) {
  return ((uint64_t)uni((uint32_t)(v >> 32)) << 32) | uni((uint32_t)v);
}
__device__ __forceinline__ double uni(double v) {
  return __builtin_bit_cast(double, uni(__builtin_bit_cast(uint64_t, v)));
}
template <class T>
__device__ __forceinline__ T* uni(T* p) {
  return (T*)uni((uint64_t)p);
}

__device__ __forceinline__ int popc64(uint64_t m) { return __popcll(m); }

// Declares that p points into workgroup memory (LDS), so the compiler
// addresses what it reaches through p with ds_ instructions at 32-bit
// addresses instead of flat ones.  Only for pointers that do point there.
#if defined(__HIP_DEVICE_COMPILE__)
#define NH_ASSUME_LDS(p) __builtin_assume(__builtin_amdgcn_is_shared((const void*)(p)))
#else
#define NH_ASSUME_LDS(p) ((void)0)
#endif

// Makes this wave's LDS writes visible to its other lanes before they read.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace nh
