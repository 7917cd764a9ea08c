// Streaming vector kernels of the Krylov solvers and smoothers, with device-resident
// scalars.  Reductions are two-stage and deterministic: each block writes one FP64 partial per
// dot product, then a single-block "finish" kernel sums the partials in a fixed order and
// evaluates the scalar recurrence (alpha, beta, ...) in FP64 on the device, so a solver
// iteration needs no host synchronisation and can be captured in a HIP graph.
#include <type_traits>

#include "common.h"
#include "blas.h"

namespace vh {

// --------------------------------------------------------------------------- reductions
// out partial[j * nblocks + block] for the dots requested by `kind`:
//   RED_DOT      x.y
//   RED_DOT2     x.y, x.w
//   RED_DOTDIFF  x.y, x.(y - w)
//   RED_NORM2    x.x
template <typename T, int KIND>
__global__ void __launch_bounds__(BLAS_BLOCK) reduce_kernel(size_t n, const T* __restrict__ x, const T* __restrict__ y,
                                                           const T* __restrict__ w, double* __restrict__ partial) {
  __shared__ double red[2][BLAS_BLOCK / 64];
  double s0 = 0.0, s1 = 0.0;
  for (size_t i = blockIdx.x * size_t(BLAS_BLOCK) + threadIdx.x; i < n; i += size_t(gridDim.x) * BLAS_BLOCK) {
    const double xi = double(x[i]);
    if (KIND == RED_NORM2) {
      s0 += xi * xi;
    } else {
      const double yi = double(y[i]);
      s0 += xi * yi;
      if (KIND == RED_DOT2) s1 += xi * double(w[i]);
      if (KIND == RED_DOTDIFF) s1 += xi * (double(y[i] - w[i]));
    }
  }
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) { red[0][wv] = s0; red[1][wv] = s1; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int k = 0; k < BLAS_BLOCK / 64; ++k) { a += red[0][k]; b += red[1][k]; }
    partial[blockIdx.x] = a;
    if (KIND == RED_DOT2 || KIND == RED_DOTDIFF) partial[gridDim.x + blockIdx.x] = b;
  }
}

// Sum `count` partial arrays of length `nparts` into scal[dst..dst+count) and evaluate `op`.
__global__ void __launch_bounds__(256) finish_kernel(const double* __restrict__ partial, int nparts, int count,
                                                    double* __restrict__ s, int dst, int op) {
  __shared__ double red[2][4];
  double acc[2] = {0.0, 0.0};
  for (int j = 0; j < count; ++j)
    for (int i = threadIdx.x; i < nparts; i += 256) acc[j] += partial[j * nparts + i];
  for (int j = 0; j < 2; ++j) acc[j] = wave_sum(acc[j]);
  if ((threadIdx.x & 63) == 0) { red[0][threadIdx.x >> 6] = acc[0]; red[1][threadIdx.x >> 6] = acc[1]; }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double v[2];
  for (int j = 0; j < 2; ++j) v[j] = red[j][0] + red[j][1] + red[j][2] + red[j][3];
  apply_scalar_op(s, v, count, dst, op);
}

// --------------------------------------------------------------------------- updates
// x += a p ; r -= a q  (a = s[S_ALPHA] * sign); optional z = dinv * r and partial r.z
template <typename T, int WITH_Z>
__global__ void __launch_bounds__(BLAS_BLOCK) cg_update_kernel(size_t n, T* __restrict__ x, const T* __restrict__ p,
                                                              T* __restrict__ r, const T* __restrict__ q,
                                                              const T* __restrict__ dinv, T* __restrict__ z,
                                                              const double* __restrict__ s, int update_r,
                                                              double* __restrict__ partial) {
  __shared__ double red[BLAS_BLOCK / 64];
  const T alpha = T(s[S_ALPHA]);
  double acc = 0.0;
  for (size_t i = blockIdx.x * size_t(BLAS_BLOCK) + threadIdx.x; i < n; i += size_t(gridDim.x) * BLAS_BLOCK) {
    x[i] = x[i] + alpha * p[i];
    if (update_r) {
      const T ri = r[i] - alpha * q[i];
      r[i] = ri;
      if (WITH_Z) {
        const T zi = dinv[i] * ri;
        z[i] = zi;
        acc += double(ri) * double(zi);
      }
    }
  }
  if (WITH_Z) {
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int k = 0; k < BLAS_BLOCK / 64; ++k) t += red[k];
      partial[blockIdx.x] = t;
    }
  }
}

// z = dinv * r, partial r.z
template <typename T>
__global__ void __launch_bounds__(BLAS_BLOCK) jacobi_dot_kernel(size_t n, const T* __restrict__ r, const T* __restrict__ dinv,
                                                               T* __restrict__ z, double* __restrict__ partial) {
  __shared__ double red[BLAS_BLOCK / 64];
  double acc = 0.0;
  for (size_t i = blockIdx.x * size_t(BLAS_BLOCK) + threadIdx.x; i < n; i += size_t(gridDim.x) * BLAS_BLOCK) {
    const T zi = dinv[i] * r[i];
    z[i] = zi;
    acc += double(r[i]) * double(zi);
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < BLAS_BLOCK / 64; ++k) t += red[k];
    partial[blockIdx.x] = t;
  }
}

// p = z + beta p   (beta = s[S_BETA]); optionally zcopy = z in the same pass
template <typename T>
__global__ void __launch_bounds__(BLAS_BLOCK) xpby_kernel(size_t n, const T* __restrict__ z, T* __restrict__ p,
                                                         const double* __restrict__ s, T* __restrict__ zcopy) {
  const T beta = T(s[S_BETA]);
  for (size_t i = blockIdx.x * size_t(BLAS_BLOCK) + threadIdx.x; i < n; i += size_t(gridDim.x) * BLAS_BLOCK) {
    const T zi = z[i];
    p[i] = zi + beta * p[i];
    if (zcopy) zcopy[i] = zi;
  }
}

// y = a * x (* w if w) + b * y   with host scalars; b == 0 never reads y.
template <typename T>
__global__ void __launch_bounds__(BLAS_BLOCK) axpby_kernel(size_t n, T a, const T* __restrict__ x, const T* __restrict__ w,
                                                          T b, T* __restrict__ y) {
  for (size_t i = blockIdx.x * size_t(BLAS_BLOCK) + threadIdx.x; i < n; i += size_t(gridDim.x) * BLAS_BLOCK) {
    T v = a * x[i];
    if (w) v *= w[i];
    y[i] = (b == T(0)) ? v : v + b * y[i];
  }
}

// Precision conversion with scaling: y = (To)(x * scale) ; or accumulate y += (To)x * scale.
template <typename Ti, typename To>
__global__ void __launch_bounds__(BLAS_BLOCK) convert_kernel(size_t n, const Ti* __restrict__ x, To* __restrict__ y,
                                                            double scale, const double* __restrict__ dscale,
                                                            int inv_scale, int accumulate) {
  double sc = dscale ? *dscale : scale;
  if (inv_scale) sc = 1.0 / sc;
  for (size_t i = blockIdx.x * size_t(BLAS_BLOCK) + threadIdx.x; i < n; i += size_t(gridDim.x) * BLAS_BLOCK) {
    if (accumulate) y[i] = To(double(y[i]) + double(x[i]) * sc);
    else y[i] = To(double(x[i]) * sc);
  }
}

// y[i] = 2 u_i - 1 in [-1, 1), u_i = (s_i >> 11) * 2^-53 with s_i the (i+1)-th output of the
// SplitMix64 generator seeded with `seed` (Steele, Lea and Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014): s_i = mix(seed + (i + 1) * 0x9E3779B97F4A7C15).  Integer
// arithmetic throughout, so any restatement reproduces every value bit for bit.
__device__ __forceinline__ unsigned long long splitmix64_at(unsigned long long seed, unsigned long long i) {
  unsigned long long z = seed + (i + 1ull) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <typename T>
__global__ void __launch_bounds__(BLAS_BLOCK) fill_splitmix_kernel(size_t n, unsigned long long seed, T* __restrict__ y) {
  for (size_t i = blockIdx.x * size_t(BLAS_BLOCK) + threadIdx.x; i < n; i += size_t(gridDim.x) * BLAS_BLOCK) {
    const double u = double(splitmix64_at(seed, i) >> 11) * 0x1.0p-53;
    y[i] = T(2.0 * u - 1.0);
  }
}

// y = bfloat16(x * scale), rounded to nearest even: BF16 storage of an FP32 vector.
__global__ void __launch_bounds__(BLAS_BLOCK) convert_bf16_kernel(size_t n, const float* __restrict__ x,
                                                                 uint16_t* __restrict__ y, double scale,
                                                                 const double* __restrict__ dscale, int inv_scale) {
  double sc = dscale ? *dscale : scale;
  if (inv_scale) sc = 1.0 / sc;
  const float s = float(sc);
  for (size_t i = blockIdx.x * size_t(BLAS_BLOCK) + threadIdx.x; i < n; i += size_t(gridDim.x) * BLAS_BLOCK)
    y[i] = f32_to_bf16_rne(s == 1.0f ? x[i] : x[i] * s);
}

// z_new = (double) z_in; partials of r.z_new and r.(z_new - z_old): the preconditioned vector
// of a flexible CG iteration taken from the V-cycle output (any precision) with both
// Polak-Ribiere dots in the same pass.
template <typename Ti, typename Tv>
__global__ void __launch_bounds__(BLAS_BLOCK) cvt_dotdiff_kernel(size_t n, const Ti* __restrict__ zin,
                                                                const Tv* __restrict__ r,
                                                                const Tv* __restrict__ zold,
                                                                Tv* __restrict__ znew,
                                                                double* __restrict__ partial) {
  __shared__ double red[2][BLAS_BLOCK / 64];
  double s0 = 0.0, s1 = 0.0;
  for (size_t i = blockIdx.x * size_t(BLAS_BLOCK) + threadIdx.x; i < n; i += size_t(gridDim.x) * BLAS_BLOCK) {
    const Tv zv = Tv(zin[i]);
    const double z = double(zv);
    const double ri = double(r[i]);
    znew[i] = zv;
    s0 += ri * z;
    s1 += ri * (z - double(zold[i]));
  }
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) { red[0][wv] = s0; red[1][wv] = s1; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int k = 0; k < BLAS_BLOCK / 64; ++k) { a += red[0][k]; b += red[1][k]; }
    partial[blockIdx.x] = a;
    partial[gridDim.x + blockIdx.x] = b;
  }
}

// Dense matrix-vector product y = mask * (A r) of the coarsest level: each wave takes
// DENSE_ROWS consecutive rows (a chunk of r a lane reads serves all of them), 4 waves per
// workgroup, and issues the loads of DENSE_UNROLL chunks before it accumulates them (one row per
// wave with eight chunks in flight measured fastest: 17.4 us for the 4,131-row FP32 inverse, 2-3
// rows per wave or fewer chunks in flight were no faster).  A is row-major with leading
// dimension ld (a multiple of 4, rows 16-byte aligned), read with 16-byte loads; the ld - n
// padding columns are never read.  Accumulation in the matrix type (FP64 for the FP64 coarsest
// inverse, also when the level vectors are FP32), one accumulator per row and lane in column
// order.
constexpr int DENSE_ROWS = 1, DENSE_UNROLL = 8;
template <typename TA, typename TV>
__global__ void __launch_bounds__(256) dense_mv_kernel(int n, int ld, const TA* __restrict__ A,
                                                      const TV* __restrict__ r, TV* __restrict__ y,
                                                      const uint8_t* __restrict__ mask) {
  constexpr int V = 16 / int(sizeof(TA));
  constexpr int R = DENSE_ROWS, UN = DENSE_UNROLL;
  using Vec = typename std::conditional<sizeof(TA) == 8, double2, float4>::type;
  const int row0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * R;
  const int lane = threadIdx.x & 63;
  if (row0 >= n) return;
  const TA* a[R];
#pragma unroll
  for (int q = 0; q < R; ++q) a[q] = A + size_t(min(row0 + q, n - 1)) * ld;   // rows past n: not stored
  const int nfull = (n / V) * V;
  TA acc[R];
#pragma unroll
  for (int q = 0; q < R; ++q) acc[q] = TA(0);
  int j = lane * V;
  for (; j + (UN - 1) * 64 * V < nfull; j += UN * 64 * V) {
    Vec x[UN][R];
    TA rv[UN][V];
#pragma unroll
    for (int u = 0; u < UN; ++u) {
#pragma unroll
      for (int q = 0; q < R; ++q) x[u][q] = *reinterpret_cast<const Vec*>(a[q] + j + u * 64 * V);
#pragma unroll
      for (int k = 0; k < V; ++k) rv[u][k] = TA(r[j + u * 64 * V + k]);
    }
#pragma unroll
    for (int u = 0; u < UN; ++u)
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const TA* e = reinterpret_cast<const TA*>(&x[u][q]);
#pragma unroll
        for (int k = 0; k < V; ++k) acc[q] += e[k] * rv[u][k];
      }
  }
  for (; j < nfull; j += 64 * V) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const Vec x = *reinterpret_cast<const Vec*>(a[q] + j);
      const TA* e = reinterpret_cast<const TA*>(&x);
#pragma unroll
      for (int k = 0; k < V; ++k) acc[q] += e[k] * TA(r[j + k]);
    }
  }
  for (int t = nfull + lane; t < n; t += 64) {
    const TA rt = TA(r[t]);
#pragma unroll
    for (int q = 0; q < R; ++q) acc[q] += a[q][t] * rt;
  }
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const TA sum = wave_sum(acc[q]);
    const int row = row0 + q;
    if (lane == 0 && row < n) y[row] = (mask && !mask[row]) ? TV(0) : TV(sum);
  }
}

// --------------------------------------------------------------------------- launchers
static inline int grid_for(size_t n) {
  size_t g = (n + BLAS_BLOCK - 1) / BLAS_BLOCK;
  if (g > BLAS_MAX_GRID) g = BLAS_MAX_GRID;
  if (g < 1) g = 1;
  return int(g);
}
int blas_grid(size_t n) { return grid_for(n); }

template <typename T>
int reduce_launch(int kind, size_t n, const T* x, const T* y, const T* w, double* partial, hipStream_t st) {
  const int g = grid_for(n);
  switch (kind) {
    case RED_DOT: hipLaunchKernelGGL((reduce_kernel<T, RED_DOT>), dim3(g), dim3(BLAS_BLOCK), 0, st, n, x, y, w, partial); break;
    case RED_DOT2: hipLaunchKernelGGL((reduce_kernel<T, RED_DOT2>), dim3(g), dim3(BLAS_BLOCK), 0, st, n, x, y, w, partial); break;
    case RED_DOTDIFF: hipLaunchKernelGGL((reduce_kernel<T, RED_DOTDIFF>), dim3(g), dim3(BLAS_BLOCK), 0, st, n, x, y, w, partial); break;
    default: hipLaunchKernelGGL((reduce_kernel<T, RED_NORM2>), dim3(g), dim3(BLAS_BLOCK), 0, st, n, x, y, w, partial); break;
  }
  return hip_status(hipGetLastError(), "reduce_kernel");
}

int finish_launch(const double* partial, int nparts, int count, double* s, int dst, int op, hipStream_t st) {
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(256), 0, st, partial, nparts, count, s, dst, op);
  return hip_status(hipGetLastError(), "finish_kernel");
}

template <typename T>
int cg_update_launch(size_t n, T* x, const T* p, T* r, const T* q, const T* dinv, T* z, const double* s,
                     int update_r, double* partial, hipStream_t st) {
  const int g = grid_for(n);
  if (dinv && update_r)
    hipLaunchKernelGGL((cg_update_kernel<T, 1>), dim3(g), dim3(BLAS_BLOCK), 0, st, n, x, p, r, q, dinv, z, s, update_r, partial);
  else
    hipLaunchKernelGGL((cg_update_kernel<T, 0>), dim3(g), dim3(BLAS_BLOCK), 0, st, n, x, p, r, q, dinv, z, s, update_r, partial);
  return hip_status(hipGetLastError(), "cg_update_kernel");
}

template <typename T>
int jacobi_dot_launch(size_t n, const T* r, const T* dinv, T* z, double* partial, hipStream_t st) {
  hipLaunchKernelGGL(jacobi_dot_kernel<T>, dim3(grid_for(n)), dim3(BLAS_BLOCK), 0, st, n, r, dinv, z, partial);
  return hip_status(hipGetLastError(), "jacobi_dot_kernel");
}

template <typename T>
int xpby_launch(size_t n, const T* z, T* p, const double* s, T* zcopy, hipStream_t st) {
  hipLaunchKernelGGL(xpby_kernel<T>, dim3(grid_for(n)), dim3(BLAS_BLOCK), 0, st, n, z, p, s, zcopy);
  return hip_status(hipGetLastError(), "xpby_kernel");
}

template <typename T>
int axpby_launch(size_t n, T a, const T* x, const T* w, T b, T* y, hipStream_t st) {
  hipLaunchKernelGGL(axpby_kernel<T>, dim3(grid_for(n)), dim3(BLAS_BLOCK), 0, st, n, a, x, w, b, y);
  return hip_status(hipGetLastError(), "axpby_kernel");
}

template <typename T>
int fill_splitmix_launch(size_t n, unsigned long long seed, T* y, hipStream_t st) {
  hipLaunchKernelGGL(fill_splitmix_kernel<T>, dim3(grid_for(n)), dim3(BLAS_BLOCK), 0, st, n, seed, y);
  return hip_status(hipGetLastError(), "fill_splitmix_kernel");
}
template int fill_splitmix_launch<double>(size_t, unsigned long long, double*, hipStream_t);
template int fill_splitmix_launch<float>(size_t, unsigned long long, float*, hipStream_t);

template <typename Ti, typename To>
int convert_launch(size_t n, const Ti* x, To* y, double scale, const double* dscale, int inv_scale, int accumulate,
                   hipStream_t st) {
  hipLaunchKernelGGL((convert_kernel<Ti, To>), dim3(grid_for(n)), dim3(BLAS_BLOCK), 0, st, n, x, y, scale, dscale,
                     inv_scale, accumulate);
  return hip_status(hipGetLastError(), "convert_kernel");
}

int convert_bf16_launch(size_t n, const float* x, uint16_t* y, double scale, const double* dscale, int inv_scale,
                        hipStream_t st) {
  hipLaunchKernelGGL(convert_bf16_kernel, dim3(grid_for(n)), dim3(BLAS_BLOCK), 0, st, n, x, y, scale, dscale, inv_scale);
  return hip_status(hipGetLastError(), "convert_bf16_kernel");
}

template <typename Ti, typename Tv>
int cvt_dotdiff_launch(size_t n, const Ti* zin, const Tv* r, const Tv* zold, Tv* znew, double* partial, hipStream_t st) {
  hipLaunchKernelGGL((cvt_dotdiff_kernel<Ti, Tv>), dim3(grid_for(n)), dim3(BLAS_BLOCK), 0, st, n, zin, r, zold, znew,
                     partial);
  return hip_status(hipGetLastError(), "cvt_dotdiff_kernel");
}
template int cvt_dotdiff_launch<double, double>(size_t, const double*, const double*, const double*, double*, double*,
                                                hipStream_t);
template int cvt_dotdiff_launch<float, double>(size_t, const float*, const double*, const double*, double*, double*,
                                               hipStream_t);
template int cvt_dotdiff_launch<float, float>(size_t, const float*, const float*, const float*, float*, double*,
                                              hipStream_t);

template <typename TA, typename TV>
int dense_mv_launch(int n, int ld, const TA* A, const TV* r, TV* y, const uint8_t* mask, hipStream_t st) {
  const int rows_per_block = 4 * DENSE_ROWS;
  hipLaunchKernelGGL((dense_mv_kernel<TA, TV>), dim3((n + rows_per_block - 1) / rows_per_block), dim3(256), 0, st, n, ld,
                     A, r, y, mask);
  return hip_status(hipGetLastError(), "dense_mv_kernel");
}
template int dense_mv_launch<double, double>(int, int, const double*, const double*, double*, const uint8_t*, hipStream_t);
template int dense_mv_launch<double, float>(int, int, const double*, const float*, float*, const uint8_t*, hipStream_t);
template int dense_mv_launch<float, float>(int, int, const float*, const float*, float*, const uint8_t*, hipStream_t);

#define VH_INST(T)                                                                                           \
  template int reduce_launch<T>(int, size_t, const T*, const T*, const T*, double*, hipStream_t);            \
  template int cg_update_launch<T>(size_t, T*, const T*, T*, const T*, const T*, T*, const double*, int, double*, hipStream_t); \
  template int jacobi_dot_launch<T>(size_t, const T*, const T*, T*, double*, hipStream_t);                    \
  template int xpby_launch<T>(size_t, const T*, T*, const double*, T*, hipStream_t);                              \
  template int axpby_launch<T>(size_t, T, const T*, const T*, T, T*, hipStream_t);
VH_INST(double)
VH_INST(float)
#undef VH_INST
template int convert_launch<double, float>(size_t, const double*, float*, double, const double*, int, int, hipStream_t);
template int convert_launch<float, double>(size_t, const float*, double*, double, const double*, int, int, hipStream_t);
template int convert_launch<double, double>(size_t, const double*, double*, double, const double*, int, int, hipStream_t);
template int convert_launch<float, float>(size_t, const float*, float*, double, const double*, int, int, hipStream_t);

}  // namespace vh
