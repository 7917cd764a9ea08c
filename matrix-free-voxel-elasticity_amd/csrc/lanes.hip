// Element-per-lanes layout control of the parity product: each element is worked by 24 lanes
// (one per node and component) of a 32-lane half of a wave64, two elements per wave.  The
// 8-point transforms are three butterfly stages over lane shuffles within the 8 lanes of one
// component, the block product gathers the three partners of each lane's block with shuffles,
// and the nodal forces are added to y with atomics (y is zeroed first; the order of the additions
// is not fixed, so results are not bitwise reproducible).  It is the reference's 24-lane layout
// (modal32, code/hex_modal_layout_v2/variants.py:95-117) on the 64-wide MI355X wave, kept as a
// control against the plane-sweep tile kernel.
#include "common.h"
#include "lanes.h"

namespace vh {

template <typename T>
struct Blocks72 {
  T c[72];
};

template <typename T>
__global__ void __launch_bounds__(256) lanes_zero_kernel(T* __restrict__ y, size_t n) {
  const size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (i < n) y[i] = T(0);
}

template <typename T>
__device__ __forceinline__ T wht8(T v, int m) {
#pragma unroll
  for (int b = 1; b < 8; b <<= 1) {
    const T other = __shfl_xor(v, b, 64);
    v = (m & b) ? other - v : v + other;
  }
  return v;
}

template <typename T>
__global__ void __launch_bounds__(256) lanes24_kernel(Blocks72<T> C, const T* __restrict__ E, const T* __restrict__ u,
                                                      T* __restrict__ y, int nx, int ny, int nz) {
  const int t = int(blockIdx.x * blockDim.x + threadIdx.x);
  const int e = t >> 5, k = t & 31, base = threadIdx.x & 32;   // element, lane in the half, half offset
  const int n_el = nx * ny * nz;
  const bool active = e < n_el && k < 24;
  const int c = active ? k >> 3 : 0, m = k & 7;                 // component, node / mode (x + 2y + 4z)
  int node = 0;
  T uv = T(0);
  if (active) {
    const int ez = e % nz, ey = (e / nz) % ny, ex = e / (nz * ny);
    node = ((ex + (m & 1)) * (ny + 1) + ey + ((m >> 1) & 1)) * (nz + 1) + ez + ((m >> 2) & 1);
    uv = u[3 * size_t(node) + c];
  }
  const T v = wht8(uv, m);
  // lane (m, c) is row c of block g = m ^ 2^c, whose columns are (mode g ^ 2^d, component d)
  const int g = m ^ (1 << c);
  T f = T(0);
#pragma unroll
  for (int d = 0; d < 3; ++d) f += C.c[9 * g + 3 * c + d] * __shfl(v, base + 8 * d + (g ^ (1 << d)), 64);
  f = wht8(f, m);
  if (active) atomicAdd(y + 3 * size_t(node) + c, E[e] * f);
}

template <typename T>
int lanes24_launch(const double* blocks72, const T* E, const T* u, T* y, int nx, int ny, int nz, hipStream_t st) {
  Blocks72<T> C;
  for (int i = 0; i < 72; ++i) C.c[i] = T(blocks72[i]);
  const size_t n = size_t(3) * (nx + 1) * (ny + 1) * (nz + 1);
  hipLaunchKernelGGL(lanes_zero_kernel<T>, dim3(int((n + 255) / 256)), dim3(256), 0, st, y, n);
  const size_t threads = size_t(nx) * ny * nz * 32;
  hipLaunchKernelGGL(lanes24_kernel<T>, dim3(int((threads + 255) / 256)), dim3(256), 0, st, C, E, u, y, nx, ny, nz);
  return hip_status(hipGetLastError(), "lanes24_kernel");
}

template int lanes24_launch<double>(const double*, const double*, const double*, double*, int, int, int, hipStream_t);
template int lanes24_launch<float>(const double*, const float*, const float*, float*, int, int, int, hipStream_t);

}  // namespace vh
