// Node- and element-owned grid kernels around the product: Jacobi diagonal, element energy
// (compliance adjoint), the dense-ke node-owned control product, multigrid transfers and
// coarsening.  All are HBM-bound; every output is written by exactly one thread (no atomics).
#include "common.h"
#include "grid_ops.h"

namespace vh {

// Local index (native node order of ke) of the corner (a, b, c) of an element, a/b/c the
// offsets along x/y/z: native order (0,0,0),(1,0,0),(1,1,0),(0,1,0),(0,0,1),(1,0,1),(1,1,1),(0,1,1).
__device__ __forceinline__ int native_of(int a, int b, int c) {
  const int lane = a + 2 * b + 4 * c;            // binary lane
  return lane ^ ((lane & 2) >> 1);               // binary <-> native is an involution
}

// ------------------------------------------------------------------ Jacobi diagonal
// diag[3n+c] = sum over incident elements of E_e ke[3i+c][3i+c] (i = the node's local index);
// with a mask, inv = 1/diag at free DOFs and 1 at fixed ones (src/voxel_mg.py:126-127).
template <typename T, int MASK>
__global__ void diag_kernel(const T* __restrict__ E, const GridArgs<T> g, T* __restrict__ out,
                            int invert) {
  const size_t n = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t nn = size_t(nx + 1) * (ny + 1) * (nz + 1);
  if (n >= nn) return;
  const int iz = int(n % (nz + 1));
  const int iy = int((n / (nz + 1)) % (ny + 1));
  const int ix = int(n / (size_t(nz + 1) * (ny + 1)));
  T acc[3] = {T(0), T(0), T(0)};
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int ex = ix - 1 + a, ey = iy - 1 + b, ez = iz - 1 + c;
        if (ex < 0 || ex >= nx || ey < 0 || ey >= ny || ez < 0 || ez >= nz) continue;
        const T Ee = E[(size_t(ex) * ny + ey) * nz + ez];
        const int li = native_of(1 - a, 1 - b, 1 - c);
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[k] += Ee * g.kediag[3 * li + k];
      }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const size_t dof = 3 * n + k;
    bool free_dof = true;
    if (MASK == MASK_X0) free_dof = ix > 0;
    if (MASK == MASK_ARRAY) free_dof = g.mask[dof] != 0;
    T v = acc[k];
    if (invert) v = free_dof ? T(1) / v : T(1);
    out[dof] = v;
  }
}

// ------------------------------------------------------------------ element energy
// ce[e] = u_e^T ke u_e through the parity blocks: with v = T u_e, u^T K u = v^T C v.
// Optional: dc[e] = -p (E0 - Emin) rho^(p-1) ce[e] (donor solver_v4.py:246-251).
template <typename T>
__global__ void energy_kernel(const T* __restrict__ u, const GridArgs<T> g, T* __restrict__ ce,
                              const T* __restrict__ rho, T* __restrict__ dc, T penal, T dE) {
  const size_t e = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t ne = size_t(nx) * ny * nz;
  if (e >= ne) return;
  const int ez = int(e % nz), ey = int((e / nz) % ny), ex = int(e / (size_t(nz) * ny));
  T v[8][3];
#pragma unroll
  for (int l = 0; l < 8; ++l) {
    const size_t node = (size_t(ex + (l & 1)) * (ny + 1) + ey + ((l >> 1) & 1)) * (nz + 1) + ez + (l >> 2);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      T x = u[3 * node + c];
      if (g.mask_kind == MASK_X0 && ex + (l & 1) == 0) x = T(0);
      if (g.mask_kind == MASK_ARRAY && !g.mask[3 * node + c]) x = T(0);
      v[l][c] = x;
    }
  }
  // in-register 8-point transform (bit 1, 2, 4 stages)
#pragma unroll
  for (int bit = 1; bit < 8; bit <<= 1)
#pragma unroll
    for (int l = 0; l < 8; ++l)
      if (!(l & bit)) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const T p = v[l][c], q = v[l | bit][c];
          v[l][c] = p + q;
          v[l | bit][c] = p - q;
        }
      }
  T s = T(0);
#pragma unroll
  for (int gi = 0; gi < 8; ++gi) {
    const T a0 = v[gi ^ 1][0], a1 = v[gi ^ 2][1], a2 = v[gi ^ 4][2];
    const T f0 = g.coef[9 * gi + 0] * a0 + g.coef[9 * gi + 1] * a1 + g.coef[9 * gi + 2] * a2;
    const T f1 = g.coef[9 * gi + 3] * a0 + g.coef[9 * gi + 4] * a1 + g.coef[9 * gi + 5] * a2;
    const T f2 = g.coef[9 * gi + 6] * a0 + g.coef[9 * gi + 7] * a1 + g.coef[9 * gi + 8] * a2;
    s += a0 * f0 + a1 * f1 + a2 * f2;
  }
  ce[e] = s;
  if (dc) {
    const T r = rho[e];
    dc[e] = -penal * dE * pow(r, penal - T(1)) * s;
  }
}

// ------------------------------------------------------------------ dense-ke control product
// Node-owned gather with the full 24x24 ke in LDS (the dense8 / node_fp64 control of the
// reference: code/hex_modal_cuda/cuda.py:47-79, donor cuda_operators.py:204-285).
template <typename T>
__global__ void __launch_bounds__(256) dense_node_kernel(const T* __restrict__ u, const T* __restrict__ E,
                                                         const T* __restrict__ ke, T* __restrict__ y,
                                                         int nx, int ny, int nz) {
  __shared__ T K[576];
  for (int i = threadIdx.x; i < 576; i += blockDim.x) K[i] = ke[i];
  __syncthreads();
  const size_t n = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  const size_t nn = size_t(nx + 1) * (ny + 1) * (nz + 1);
  if (n >= nn) return;
  const int iz = int(n % (nz + 1));
  const int iy = int((n / (nz + 1)) % (ny + 1));
  const int ix = int(n / (size_t(nz + 1) * (ny + 1)));
  T acc[3] = {T(0), T(0), T(0)};
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) {
        const int ex = ix - 1 + a, ey = iy - 1 + b, ez = iz - 1 + c;
        if (ex < 0 || ex >= nx || ey < 0 || ey >= ny || ez < 0 || ez >= nz) continue;
        const T Ee = E[(size_t(ex) * ny + ey) * nz + ez];
        const int li = native_of(1 - a, 1 - b, 1 - c);
        T loc[3] = {T(0), T(0), T(0)};
        for (int m = 0; m < 8; ++m) {
          const int dx = (m == 1 || m == 2 || m == 5 || m == 6);
          const int dy = (m == 2 || m == 3 || m == 6 || m == 7);
          const int dz = (m >= 4);
          const size_t node = (size_t(ex + dx) * (ny + 1) + ey + dy) * (nz + 1) + ez + dz;
          const T u0 = u[3 * node], u1 = u[3 * node + 1], u2 = u[3 * node + 2];
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const T* row = K + (3 * li + k) * 24 + 3 * m;
            loc[k] += row[0] * u0 + row[1] * u1 + row[2] * u2;
          }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[k] += Ee * loc[k];
      }
#pragma unroll
  for (int k = 0; k < 3; ++k) y[3 * n + k] = acc[k];
}

// ------------------------------------------------------------------ transfers
// One workgroup per z-line (restrict_line / prolong_line in grid_ops.h).
template <typename T>
__global__ void __launch_bounds__(256) restrict_kernel(const T* __restrict__ rf, T* __restrict__ rc,
                                                      const GridArgs<T> g, const uint8_t* __restrict__ cmask,
                                                      int cmask_kind) {
  extern __shared__ unsigned char smem_raw[];
  restrict_line<T>(rf, rc, g.nx, g.ny, g.nz, g.mask_kind, g.mask, cmask, cmask_kind, int(blockIdx.x),
                   reinterpret_cast<T*>(smem_raw));
}

// prolong: xf (+)= mask_f * P ec.
template <typename T>
__global__ void __launch_bounds__(256) prolong_kernel(const T* __restrict__ ec, T* __restrict__ xf,
                                                     const GridArgs<T> g, int assign) {
  extern __shared__ unsigned char smem_raw[];
  prolong_line<T>(ec, xf, g.nx, g.ny, g.nz, g.mask_kind, g.mask, assign, int(blockIdx.x),
                  reinterpret_cast<T*>(smem_raw));
}

// 2x2x2 mean of element moduli (src/voxel_mg.py:108-111); g = fine grid.
template <typename T>
__global__ void coarsen_modulus_kernel(const T* __restrict__ Ef, T* __restrict__ Ec, int nx, int ny, int nz) {
  const int cx = nx / 2, cy = ny / 2, cz = nz / 2;
  const size_t e = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (e >= size_t(cx) * cy * cz) return;
  const int ez = int(e % cz), ey = int((e / cz) % cy), ex = int(e / (size_t(cz) * cy));
  // same summation order as a reshape(.., 2, .., 2, .., 2).mean over axes (1, 3, 5)
  T s = T(0);
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c)
        s += Ef[(size_t(2 * ex + a) * ny + 2 * ey + b) * nz + 2 * ez + c];
  Ec[e] = s / T(8);
}

// ------------------------------------------------------------------ launchers
static inline int blocks_for(size_t n, int bs) { return int((n + bs - 1) / bs); }

template <typename T>
int diag_launch(const T* E, const GridArgs<T>& g, T* out, int invert, hipStream_t st) {
  const size_t nn = size_t(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  const int bs = 256;
  switch (g.mask_kind) {
    case MASK_NONE: hipLaunchKernelGGL((diag_kernel<T, MASK_NONE>), dim3(blocks_for(nn, bs)), dim3(bs), 0, st, E, g, out, invert); break;
    case MASK_ARRAY: hipLaunchKernelGGL((diag_kernel<T, MASK_ARRAY>), dim3(blocks_for(nn, bs)), dim3(bs), 0, st, E, g, out, invert); break;
    default: hipLaunchKernelGGL((diag_kernel<T, MASK_X0>), dim3(blocks_for(nn, bs)), dim3(bs), 0, st, E, g, out, invert); break;
  }
  return hip_status(hipGetLastError(), "diag_kernel");
}

template <typename T>
int energy_launch(const T* u, const GridArgs<T>& g, T* ce, const T* rho, T* dc, T penal, T dE, hipStream_t st) {
  const size_t ne = size_t(g.nx) * g.ny * g.nz;
  hipLaunchKernelGGL(energy_kernel<T>, dim3(blocks_for(ne, 256)), dim3(256), 0, st, u, g, ce, rho, dc, penal, dE);
  return hip_status(hipGetLastError(), "energy_kernel");
}

template <typename T>
int dense_launch(const T* u, const T* E, const T* ke, T* y, int nx, int ny, int nz, hipStream_t st) {
  const size_t nn = size_t(nx + 1) * (ny + 1) * (nz + 1);
  hipLaunchKernelGGL(dense_node_kernel<T>, dim3(blocks_for(nn, 256)), dim3(256), 0, st, u, E, ke, y, nx, ny, nz);
  return hip_status(hipGetLastError(), "dense_node_kernel");
}

template <typename T>
int restrict_launch(const T* rf, T* rc, const GridArgs<T>& g, const uint8_t* cmask, int cmask_kind, hipStream_t st) {
  const unsigned lines = unsigned(g.nx / 2 + 1) * unsigned(g.ny / 2 + 1);
  const size_t lds = size_t(3) * (g.nz + 1) * sizeof(T);
  hipLaunchKernelGGL(restrict_kernel<T>, dim3(lines), dim3(256), lds, st, rf, rc, g, cmask, cmask_kind);
  return hip_status(hipGetLastError(), "restrict_kernel");
}

template <typename T>
int prolong_launch(const T* ec, T* xf, const GridArgs<T>& g, int assign, hipStream_t st) {
  const unsigned lines = unsigned(g.nx + 1) * unsigned(g.ny + 1);
  const size_t lds = size_t(3) * (g.nz / 2 + 1) * sizeof(T);
  hipLaunchKernelGGL(prolong_kernel<T>, dim3(lines), dim3(256), lds, st, ec, xf, g, assign);
  return hip_status(hipGetLastError(), "prolong_kernel");
}

template <typename T>
int coarsen_modulus_launch(const T* Ef, T* Ec, int nx, int ny, int nz, hipStream_t st) {
  const size_t ne = size_t(nx / 2) * (ny / 2) * (nz / 2);
  hipLaunchKernelGGL(coarsen_modulus_kernel<T>, dim3(blocks_for(ne, 256)), dim3(256), 0, st, Ef, Ec, nx, ny, nz);
  return hip_status(hipGetLastError(), "coarsen_modulus_kernel");
}

#define VH_INST(T)                                                                              \
  template int diag_launch<T>(const T*, const GridArgs<T>&, T*, int, hipStream_t);              \
  template int energy_launch<T>(const T*, const GridArgs<T>&, T*, const T*, T*, T, T, hipStream_t); \
  template int dense_launch<T>(const T*, const T*, const T*, T*, int, int, int, hipStream_t);    \
  template int restrict_launch<T>(const T*, T*, const GridArgs<T>&, const uint8_t*, int, hipStream_t); \
  template int prolong_launch<T>(const T*, T*, const GridArgs<T>&, int, hipStream_t);           \
  template int coarsen_modulus_launch<T>(const T*, T*, int, int, int, hipStream_t);
VH_INST(double)
VH_INST(float)
#undef VH_INST

}  // namespace vh
