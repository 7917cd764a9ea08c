// Shared definitions of the voxelhip HIP library (gfx950 / CDNA4, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace vh {

// Error state of the C-ABI: the last failing call's message, per thread.
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);
int hip_status(hipError_t e, const char* what);

// Storage conversions.  bf16 is stored as its upper 16 bits of an IEEE binary32.
__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
  uint32_t x = __float_as_uint(f);
  uint32_t lsb = (x >> 16) & 1u;
  x += 0x7fffu + lsb;                // round to nearest even
  return uint16_t(x >> 16);
}

template <typename S, typename T> __device__ __forceinline__ T load_as(const S* p, size_t i);
template <> __device__ __forceinline__ double load_as<double, double>(const double* p, size_t i) { return p[i]; }
template <> __device__ __forceinline__ float load_as<float, float>(const float* p, size_t i) { return p[i]; }
template <> __device__ __forceinline__ double load_as<float, double>(const float* p, size_t i) { return double(p[i]); }
template <> __device__ __forceinline__ float load_as<uint16_t, float>(const uint16_t* p, size_t i) {
  return bf16_to_f32(p[i]);
}

// Wave64 sum with a fixed butterfly order (deterministic).
template <typename T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum of one value per thread; result valid in thread 0.  `red` holds >= nwaves entries.
template <typename T, int BLOCK> __device__ __forceinline__ T block_sum(T v, T* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  T s = T(0);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < (BLOCK + 63) / 64; ++i) s += red[i];
  }
  return s;
}

// XCD-aware work decode for grids that are multiples of 8: the hardware assigns block b to XCD
// b % 8; XCD k gets the contiguous logical range [k*per, (k+1)*per), so neighbouring work units
// (shared halo rows, columns and planes) are mostly served by the same L2.
__device__ __forceinline__ int xcd_logical(int b, int grid) { return (b & 7) * (grid >> 3) + (b >> 3); }

// Dtype codes shared with the Python side.
enum Dtype : int { F64 = 0, F32 = 1, BF16 = 2 };
// Mask kinds: no mask (raw product), a per-DOF uint8 array (1 = free), the x = 0 clamp.
enum MaskKind : int { MASK_NONE = 0, MASK_ARRAY = 1, MASK_X0 = 2 };

}  // namespace vh
