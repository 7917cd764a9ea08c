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
// One workgroup per block (restrict_block / prolong_block in grid_ops.h), XCD-aware order so
// that the blocks sharing fine planes (restriction) or coarse rows (prolongation) meet in one L2.
template <typename T>
__global__ void __launch_bounds__(TR_RESTRICT::NT) restrict_kernel(const T* __restrict__ rf, T* __restrict__ rc,
                                                             const GridArgs<T> g, const uint8_t* __restrict__ cmask,
                                                             int cmask_kind, int rows, int nblocks) {
  extern __shared__ unsigned char smem_raw[];
  const int blk = xcd_logical(blockIdx.x, gridDim.x);
  if (blk >= nblocks) return;
  restrict_block<T>(rf, rc, g.nx, g.ny, g.nz, g.mask_kind, g.mask, cmask, cmask_kind, rows, blk,
                    reinterpret_cast<T*>(smem_raw));
}

// prolong: xf (+)= mask_f * P ec.
template <typename T>
__global__ void __launch_bounds__(TR_PROLONG::NT) prolong_kernel(const T* __restrict__ ec, T* __restrict__ xf,
                                                            const GridArgs<T> g, int assign, int rows, int nblocks) {
  extern __shared__ unsigned char smem_raw[];
  const int blk = xcd_logical(blockIdx.x, gridDim.x);
  if (blk >= nblocks) return;
  prolong_block<T>(ec, xf, g.nx, g.ny, g.nz, g.mask_kind, g.mask, assign, rows, blk,
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

// ------------------------------------------------------------------ dense coarsest operators
// Free flag of DOF i under a mask (X0: every DOF of the plane ix = 0 is fixed).
__device__ __forceinline__ bool dof_free(int mask_kind, const uint8_t* mask, int i, int plane_dofs) {
  if (mask_kind == MASK_ARRAY) return mask[i] != 0;
  if (mask_kind == MASK_X0) return i >= plane_dofs;
  return true;
}

// A = M K M + (I - M) of a small grid, one thread per entry (i, j) of the n x n matrix (row
// stride ld): K_ij sums E_e ke[(I in e, a), (J in e, b)] over the elements e shared by nodes I
// and J, in ascending element order (src/voxel_mg.py:61-70).
__global__ void __launch_bounds__(256) assemble_dense_kernel(const double* __restrict__ ke,
                                                             const double* __restrict__ E, int nx, int ny, int nz,
                                                             int mask_kind, const uint8_t* __restrict__ mask,
                                                             double* __restrict__ A, int ld) {
#pragma clang fp contract(off)   // E_e * ke rounded before the sum, as an assembly on the host does
  const int ny1 = ny + 1, nz1 = nz + 1;
  const int n = 3 * (nx + 1) * ny1 * nz1;
  const size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (t >= size_t(n) * n) return;
  const int i = int(t / unsigned(n)), j = int(t - size_t(i) * n);
  const int I = i / 3, a = i - 3 * I, J = j / 3, b = j - 3 * J;
  const int iz = I % nz1, iy = (I / nz1) % ny1, ix = I / (nz1 * ny1);
  const int jz = J % nz1, jy = (J / nz1) % ny1, jx = J / (nz1 * ny1);
  double k = 0.0;
  if (abs(ix - jx) <= 1 && abs(iy - jy) <= 1 && abs(iz - jz) <= 1) {
    for (int ex = max(max(ix, jx) - 1, 0); ex <= min(min(ix, jx), nx - 1); ++ex)
      for (int ey = max(max(iy, jy) - 1, 0); ey <= min(min(iy, jy), ny - 1); ++ey)
        for (int ez = max(max(iz, jz) - 1, 0); ez <= min(min(iz, jz), nz - 1); ++ez) {
          const int li = native_of(ix - ex, iy - ey, iz - ez), lj = native_of(jx - ex, jy - ey, jz - ez);
          k += E[(size_t(ex) * ny + ey) * nz + ez] * ke[(3 * li + a) * 24 + 3 * lj + b];
        }
  }
  const int plane = 3 * ny1 * nz1;
  const bool fi = dof_free(mask_kind, mask, i, plane), fj = dof_free(mask_kind, mask, j, plane);
  A[size_t(i) * ld + j] = (fi && fj) ? k : ((i == j && !fi) ? 1.0 : 0.0);
}

// Galerkin coarse operator from 81 probe products: row p = ((px * 3 + py) * 3 + pz) * 3 + c of G
// holds R A_f P v_p for the probe v_p = sum of e_(J, c) over the coarse nodes J = (px, py, pz)
// mod 3.  P^T A_f P couples only nodes within one node of each other, so entry (I, a) of row p
// belongs to the column of the one J = p's residues within one node of I.  Writes
// M_c A_c M_c + (I - M_c) into those entries (A zeroed beforehand: every other entry is 0).
__global__ void __launch_bounds__(256) galerkin_scatter_kernel(const double* __restrict__ G, int nx, int ny, int nz,
                                                               int mask_kind, const uint8_t* __restrict__ mask,
                                                               double* __restrict__ A, int ld) {
  const int nx1 = nx + 1, ny1 = ny + 1, nz1 = nz + 1;
  const int n = 3 * nx1 * ny1 * nz1;
  const size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (t >= size_t(81) * n) return;
  const int p = int(t / unsigned(n)), i = int(t - size_t(p) * n);
  const int c = p % 3, pz = (p / 3) % 3, py = (p / 9) % 3, px = p / 27;
  const int I = i / 3;
  const int iz = I % nz1, iy = (I / nz1) % ny1, ix = I / (nz1 * ny1);
  // the J in {I - 1, I, I + 1} with J = residue (mod 3), per axis
  const int jx = ix - 1 + (px - ix + 1 + 3 * nx1) % 3;
  const int jy = iy - 1 + (py - iy + 1 + 3 * ny1) % 3;
  const int jz = iz - 1 + (pz - iz + 1 + 3 * nz1) % 3;
  if (jx < 0 || jx >= nx1 || jy < 0 || jy >= ny1 || jz < 0 || jz >= nz1) return;
  const int j = 3 * ((jx * ny1 + jy) * nz1 + jz) + c;
  const int plane = 3 * ny1 * nz1;
  const bool fi = dof_free(mask_kind, mask, i, plane), fj = dof_free(mask_kind, mask, j, plane);
  A[size_t(i) * ld + j] = (fi && fj) ? G[size_t(p) * n + i] : ((i == j && !fi) ? 1.0 : 0.0);
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

// A single-problem transfer keeps at least two blocks per CU (fewer, fuller blocks would leave a
// small grid's launch latency-bound on fewer CUs); the CU count is read once.
static int transfer_min_blocks() {
  static const int n = [] {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
    return 2 * cus;
  }();
  return n;
}

template <typename T>
int restrict_launch(const T* rf, T* rc, const GridArgs<T>& g, const uint8_t* cmask, int cmask_kind, hipStream_t st) {
  const int rows = restrict_rows<TR_RESTRICT>(g.nx, g.ny, g.nz, transfer_min_blocks());
  const TransferGeom tg = restrict_geom<TR_RESTRICT>(g.nx, g.ny, g.nz, rows);
  const unsigned grid = unsigned((tg.nblocks + 7) / 8) * 8;
  hipLaunchKernelGGL(restrict_kernel<T>, dim3(grid), dim3(TR_RESTRICT::NT), tg.lds * sizeof(T), st, rf, rc, g, cmask,
                     cmask_kind, rows, tg.nblocks);
  return hip_status(hipGetLastError(), "restrict_kernel");
}

template <typename T>
int prolong_launch(const T* ec, T* xf, const GridArgs<T>& g, int assign, hipStream_t st) {
  const int rows = prolong_rows<TR_PROLONG>(g.nx, g.ny, g.nz, transfer_min_blocks());
  const TransferGeom tg = prolong_geom<TR_PROLONG>(g.nx, g.ny, g.nz, rows);
  const unsigned grid = unsigned((tg.nblocks + 7) / 8) * 8;
  hipLaunchKernelGGL(prolong_kernel<T>, dim3(grid), dim3(TR_PROLONG::NT), tg.lds * sizeof(T), st, ec, xf, g, assign,
                     rows, tg.nblocks);
  return hip_status(hipGetLastError(), "prolong_kernel");
}

int assemble_dense_launch(const double* ke, const double* E, int nx, int ny, int nz, int mask_kind,
                          const uint8_t* mask, double* A, int ld, hipStream_t st) {
  const size_t n = size_t(3) * (nx + 1) * (ny + 1) * (nz + 1);
  hipLaunchKernelGGL(assemble_dense_kernel, dim3(blocks_for(n * n, 256)), dim3(256), 0, st, ke, E, nx, ny, nz,
                     mask_kind, mask, A, ld);
  return hip_status(hipGetLastError(), "assemble_dense_kernel");
}

int galerkin_scatter_launch(const double* G, int nx, int ny, int nz, int mask_kind, const uint8_t* mask, double* A,
                            int ld, hipStream_t st) {
  const size_t n = size_t(3) * (nx + 1) * (ny + 1) * (nz + 1);
  hipLaunchKernelGGL(galerkin_scatter_kernel, dim3(blocks_for(81 * n, 256)), dim3(256), 0, st, G, nx, ny, nz,
                     mask_kind, mask, A, ld);
  return hip_status(hipGetLastError(), "galerkin_scatter_kernel");
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
