// Assembled coarse operators of the Galerkin hierarchy: symmetric 27-point node-block stencils.
//
// A coarse operator K = P^T A P of trilinear P couples each node I only with the nodes I + d,
// d in {-1, 0, 1}^3 (27 neighbours).  Offset index k = ((dx+1)*3 + (dy+1))*3 + (dz+1), the centre
// is k = 13, and -d has index 26 - k.  The stencil stores the 3x3 blocks K[I, I + d_k] of the
// upper half k = 13..26 as slot j = k - 13 (14 slots, 126 values per node), coefficient-major so
// that a wave reading one coefficient of 64 consecutive nodes touches 64 consecutive words:
//     U[((j * 3 + r) * 3 + c) * nn + I] = K[3 I + r, 3 (I + d_(13+j)) + c]   (0 outside the grid)
// The lower half is the transpose of the neighbour's slot: K[I, I + d_k] = K[I + d_k, I]^T with
// K[I + d_k, I] held by node I + d_k in slot 13 - k.
#pragma once
#include "common.h"

namespace vh {

template <typename T>
struct StencilParams {
  const T* U;              // 14 x 9 x nn coefficients
  const T* u;              // input (nodal, 3 per node)
  T* y;                    // output (never aliases u)
  const T* b;              // right-hand side (RESID, JACOBI, CHEB3)
  const T* dinv;           // inverse diagonal (JACOBI, CHEB3)
  const T* prev;           // CHEB3: previous iterate (default: y on entry)
  float* y32;              // RESID: FP32 copy of the output
  const uint8_t* mask;     // per-DOF free flags (MASK_ARRAY)
  int nx, ny, nz;          // element grid (nodes: (nx+1) (ny+1) (nz+1))
  int fixed_zero, flags;
  T omega, cheb_a, cheb_c;
};

template <typename T>
int stencil_apply_launch(const StencilParams<T>& p, int mode, int mask_kind, hipStream_t st);
template <typename T>
int stencil_scatter_launch(const double* G, T* U, int nx, int ny, int nz, hipStream_t st);

}  // namespace vh
