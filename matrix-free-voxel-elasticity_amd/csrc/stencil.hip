// Galerkin coarse operators as 27-point node-block stencils: assembly from probe products and
// the node-owned product with the fused epilogues of the V-cycle's coarse-level passes.
//
// The donor Galerkin hierarchy (donor_yang_gmg multigrid_v4.py:132-270,747-772) assembles
// R K P into CSR matrices with host-built sparsity and applies them with cuSPARSE.  Here the
// operator of a voxel grid is a fixed 27-point pattern, so it is stored as dense per-node
// blocks (no indices) and applied by one thread per node: the thread reads the 27 neighbour
// nodes' inputs and the 243 block coefficients (126 of its own, 117 transposed from the
// neighbours that hold the lower half), both coalesced across a wave, and sums them in a
// fixed order (bitwise reproducible).
#include "matvec.h"
#include "stencil.h"

namespace vh {

__device__ __forceinline__ bool node_dof_free(int mask_kind, const uint8_t* mask, int dof, int plane_dofs) {
  if (mask_kind == MASK_ARRAY) return mask[dof] != 0;
  if (mask_kind == MASK_X0) return dof >= plane_dofs;
  return true;
}

// U from the 81 probe products G (row p = 3 * class(J) + c holds R A P v_p for the probe summing
// e_(J, c) over the nodes J of residue class (Jx, Jy, Jz) mod 3): K[3I + r, 3J + c] is the entry
// 3I + r of the probe of J's class, J being the only node of that class next to I.
template <typename T>
__global__ void __launch_bounds__(256) stencil_scatter_kernel(const double* __restrict__ G, T* __restrict__ U, int nx,
                                                              int ny, int nz) {
  const int nx1 = nx + 1, ny1 = ny + 1, nz1 = nz + 1;
  const int nn = nx1 * ny1 * nz1;
  const size_t t = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (t >= size_t(126) * nn) return;
  const int q = int(t / unsigned(nn)), I = int(t - size_t(q) * nn);
  const int j = q / 9, r = (q / 3) % 3, c = q % 3;
  const int k = 13 + j;
  const int dx = k / 9 - 1, dy = (k / 3) % 3 - 1, dz = k % 3 - 1;
  const int iz = I % nz1, iy = (I / nz1) % ny1, ix = I / (nz1 * ny1);
  const int jx = ix + dx, jy = iy + dy, jz = iz + dz;
  double v = 0.0;
  if (jx >= 0 && jx < nx1 && jy >= 0 && jy < ny1 && jz >= 0 && jz < nz1) {
    const int cls = ((jx % 3) * 3 + jy % 3) * 3 + jz % 3;
    v = G[size_t(3 * cls + c) * (3 * size_t(nn)) + 3 * size_t(I) + r];
  }
  U[t] = T(v);
}

// One thread per node: A u = M K M u + (I - M) u (K the stencil), then the epilogue of MODE.
template <typename T, int MODE, int MASK>
__global__ void __launch_bounds__(256) stencil_apply_kernel(StencilParams<T> a) {
  const int nx1 = a.nx + 1, ny1 = a.ny + 1, nz1 = a.nz + 1;
  const int nn = nx1 * ny1 * nz1;
  const int I = blockIdx.x * 256 + threadIdx.x;
  if (I >= nn) return;
  const int iz = I % nz1, iy = (I / nz1) % ny1, ix = I / (nz1 * ny1);
  const int plane_dofs = 3 * ny1 * nz1;
  const T* __restrict__ U = a.U;
  const T* __restrict__ u = a.u;
  T acc[3] = {T(0), T(0), T(0)};
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    const int dx = k / 9 - 1, dy = (k / 3) % 3 - 1, dz = k % 3 - 1;
    const int jx = ix + dx, jy = iy + dy, jz = iz + dz;
    if (jx < 0 || jx >= nx1 || jy < 0 || jy >= ny1 || jz < 0 || jz >= nz1) continue;
    const int J = (jx * ny1 + jy) * nz1 + jz;
    T uj[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      uj[c] = u[3 * size_t(J) + c];
      if (MASK == MASK_X0 && jx == 0) uj[c] = T(0);
      if (MASK == MASK_ARRAY && a.mask[3 * size_t(J) + c] == 0) uj[c] = T(0);
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      T s = acc[r];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        // upper half: own slot k - 13; lower half: the neighbour's slot 13 - k, transposed
        const T kv = k >= 13 ? U[size_t(((k - 13) * 3 + r) * 3 + c) * nn + I]
                             : U[size_t(((13 - k) * 3 + c) * 3 + r) * nn + J];
        s += kv * uj[c];
      }
      acc[r] = s;
    }
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const size_t dof = 3 * size_t(I) + r;
    const bool free_dof = node_dof_free(MASK, a.mask, int(dof), plane_dofs);
    const T uo = u[dof];
    const T au = (MASK == MASK_NONE || free_dof) ? acc[r] : (a.fixed_zero ? T(0) : uo);
    T out;
    if (MODE == MODE_RAW) {
      out = au;
    } else if (MODE == MODE_RESID) {
      out = a.b[dof] - au;
    } else {
      const T z = a.dinv[dof] * (a.b[dof] - au);
      if (MODE == MODE_JACOBI) {
        out = uo + a.omega * z;
      } else {   // CHEB3
        const T pv = (a.cheb_c == T(0) || (a.flags & MV_PREV_ZERO)) ? T(0) : (a.prev ? a.prev : a.y)[dof];
        out = uo + a.cheb_a * (a.cheb_c == T(0) ? z : z + a.cheb_c * (uo - pv));
      }
    }
    if (MASK != MASK_NONE && a.fixed_zero && !free_dof) out = T(0);
    if (a.y) a.y[dof] = out;
    if (MODE == MODE_RESID && a.y32) a.y32[dof] = float(out);
  }
}

static inline int nblocks(size_t n, int bs) { return int((n + bs - 1) / bs); }

template <typename T, int MODE>
static void apply_mask(const StencilParams<T>& p, int mask_kind, int grid, hipStream_t st) {
  switch (mask_kind) {
    case MASK_NONE: hipLaunchKernelGGL((stencil_apply_kernel<T, MODE, MASK_NONE>), dim3(grid), dim3(256), 0, st, p); break;
    case MASK_ARRAY: hipLaunchKernelGGL((stencil_apply_kernel<T, MODE, MASK_ARRAY>), dim3(grid), dim3(256), 0, st, p); break;
    default: hipLaunchKernelGGL((stencil_apply_kernel<T, MODE, MASK_X0>), dim3(grid), dim3(256), 0, st, p); break;
  }
}

template <typename T>
int stencil_apply_launch(const StencilParams<T>& p, int mode, int mask_kind, hipStream_t st) {
  const size_t nn = size_t(p.nx + 1) * (p.ny + 1) * (p.nz + 1);
  const int grid = nblocks(nn, 256);
  switch (mode) {
    case MODE_RAW: apply_mask<T, MODE_RAW>(p, mask_kind, grid, st); break;
    case MODE_RESID: apply_mask<T, MODE_RESID>(p, mask_kind, grid, st); break;
    case MODE_JACOBI: apply_mask<T, MODE_JACOBI>(p, mask_kind, grid, st); break;
    case MODE_CHEB3: apply_mask<T, MODE_CHEB3>(p, mask_kind, grid, st); break;
    default: return fail(-2, "stencil product: mode %d not available (RAW, RESID, JACOBI, CHEB3)", mode);
  }
  return hip_status(hipGetLastError(), "stencil_apply_kernel");
}

template <typename T>
int stencil_scatter_launch(const double* G, T* U, int nx, int ny, int nz, hipStream_t st) {
  const size_t nn = size_t(nx + 1) * (ny + 1) * (nz + 1);
  hipLaunchKernelGGL((stencil_scatter_kernel<T>), dim3(nblocks(126 * nn, 256)), dim3(256), 0, st, G, U, nx, ny, nz);
  return hip_status(hipGetLastError(), "stencil_scatter_kernel");
}

template int stencil_apply_launch<double>(const StencilParams<double>&, int, int, hipStream_t);
template int stencil_apply_launch<float>(const StencilParams<float>&, int, int, hipStream_t);
template int stencil_scatter_launch<double>(const double*, double*, int, int, int, hipStream_t);
template int stencil_scatter_launch<float>(const double*, float*, int, int, int, hipStream_t);

}  // namespace vh
