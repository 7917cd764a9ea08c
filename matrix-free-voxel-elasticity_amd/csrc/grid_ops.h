#pragma once
#include "common.h"

namespace vh {

template <typename T>
struct GridArgs {
  int nx, ny, nz;           // element grid
  int mask_kind;            // MaskKind of `mask`
  const uint8_t* mask;      // per-DOF free flags (MASK_ARRAY)
  T kediag[24];             // diagonal of ke (native order)
  T coef[72];               // parity blocks
};


// ------------------------------------------------------------------ transfers (device parts)
// Trilinear weights along one axis: fine index i against coarse index I: w = 1 if i == 2I,
// 1/2 if |i - 2I| == 1 (src/voxel_mg.py:74-105: P copies even nodes, averages odd ones; R = P^T).
// One workgroup handles one z-line of nodes (all iz, 3 components: a contiguous run of
// 3 (nz + 1) values in the author layout), so every global access is a coalesced sweep along a
// line.  restrict: the x/y weights combine up to 9 fine lines into one weighted fine line in
// LDS (`line`, 3 (nz + 1) entries), then each coarse DOF takes its 3 z neighbours from LDS.
// prolong: up to 4 coarse lines are combined with the x/y weights into LDS (3 (nz/2 + 1)
// entries), then each fine DOF interpolates along z.  (nx, ny, nz) is the FINE element grid.

// rc[I] = mask_c[I] * sum_{a,b,c in -1..1} w(a) w(b) w(c) (rf * mask_f)[2I + (a,b,c)] on coarse
// line `L` = Ix * (ny/2 + 1) + Iy.
template <typename T>
__device__ __forceinline__ void restrict_line(const T* __restrict__ rf, T* __restrict__ rc, int nx, int ny, int nz,
                                              int fmask_kind, const uint8_t* __restrict__ fmask,
                                              const uint8_t* __restrict__ cmask, int cmask_kind, int L, T* line) {
  const int cy1 = ny / 2 + 1, cz1 = nz / 2 + 1;
  const int fy1 = ny + 1, nlf = 3 * (nz + 1), nlc = 3 * cz1;
  const int Iy = int(unsigned(L) % unsigned(cy1)), Ix = int(unsigned(L) / unsigned(cy1));
  for (int t = threadIdx.x; t < nlf; t += blockDim.x) {
    T acc = T(0);
#pragma unroll
    for (int a = -1; a <= 1; ++a) {
      const int ix = 2 * Ix + a;
      if (ix < 0 || ix > nx || (fmask_kind == MASK_X0 && ix == 0)) continue;
      const T wa = a == 0 ? T(1) : T(0.5);
#pragma unroll
      for (int b = -1; b <= 1; ++b) {
        const int iy = 2 * Iy + b;
        if (iy < 0 || iy > ny) continue;
        const size_t dof = (size_t(ix) * fy1 + iy) * nlf + t;
        T v = rf[dof];
        if (fmask_kind == MASK_ARRAY && !fmask[dof]) v = T(0);
        acc += (wa * (b == 0 ? T(1) : T(0.5))) * v;
      }
    }
    line[t] = acc;
  }
  __syncthreads();
  const size_t cbase = (size_t(Ix) * cy1 + Iy) * nlc;
  for (int t = threadIdx.x; t < nlc; t += blockDim.x) {
    const int Iz = t / 3, k = t - 3 * Iz;
    T acc = line[6 * Iz + k];
    if (Iz > 0) acc += T(0.5) * line[6 * Iz - 3 + k];
    if (2 * Iz + 1 <= nz) acc += T(0.5) * line[6 * Iz + 3 + k];
    bool free_dof = true;
    if (cmask_kind == MASK_X0) free_dof = Ix > 0;
    if (cmask_kind == MASK_ARRAY) free_dof = cmask[cbase + t] != 0;
    rc[cbase + t] = free_dof ? acc : T(0);
  }
}

// xf (+)= mask_f * P ec on fine line `L` = ix * (ny + 1) + iy.
template <typename T>
__device__ __forceinline__ void prolong_line(const T* __restrict__ ec, T* __restrict__ xf, int nx, int ny, int nz,
                                             int mask_kind, const uint8_t* __restrict__ mask, int assign, int L,
                                             T* cl) {
  const int fy1 = ny + 1, nlf = 3 * (nz + 1);
  const int cy1 = ny / 2 + 1, nlc = 3 * (nz / 2 + 1);
  const int iy = int(unsigned(L) % unsigned(fy1)), ix = int(unsigned(L) / unsigned(fy1));
  const int x0 = ix >> 1, x1 = (ix + 1) >> 1, y0 = iy >> 1, y1 = (iy + 1) >> 1;
  const T wxy = ((ix & 1) ? T(0.5) : T(1)) * ((iy & 1) ? T(0.5) : T(1));
  const T* c00 = ec + (size_t(x0) * cy1 + y0) * nlc;
  const T* c01 = ec + (size_t(x0) * cy1 + y1) * nlc;
  const T* c10 = ec + (size_t(x1) * cy1 + y0) * nlc;
  const T* c11 = ec + (size_t(x1) * cy1 + y1) * nlc;
  const size_t fbase = (size_t(ix) * fy1 + iy) * nlf;
  // the fine values this thread updates are loaded before the coarse lines, so both memory
  // round trips of the workgroup overlap (lines of up to 3 * 2049 values: <= 24 per thread)
  constexpr int PF = 4;
  T xold[PF];
  if (!assign) {
#pragma unroll
    for (int j = 0; j < PF; ++j) {
      const int t = threadIdx.x + j * blockDim.x;
      xold[j] = t < nlf ? xf[fbase + t] : T(0);
    }
  }
  for (int t = threadIdx.x; t < nlc; t += blockDim.x) {
    T v = c00[t];
    if (iy & 1) v += c01[t];
    if (ix & 1) {
      v += c10[t];
      if (iy & 1) v += c11[t];
    }
    cl[t] = wxy * v;
  }
  __syncthreads();
  const bool plane_free = !(mask_kind == MASK_X0 && ix == 0);
  auto interp = [&](int t) {
    const int iz = t / 3, k = t - 3 * iz, z0 = iz >> 1;
    T p = cl[3 * z0 + k];
    if (iz & 1) p = T(0.5) * (p + cl[3 * z0 + 3 + k]);
    bool free_dof = plane_free;
    if (mask_kind == MASK_ARRAY) free_dof = mask[fbase + t] != 0;
    return free_dof ? p : T(0);
  };
#pragma unroll
  for (int j = 0; j < PF; ++j) {
    const int t = threadIdx.x + j * blockDim.x;
    if (t < nlf) xf[fbase + t] = (assign ? T(0) : xold[j]) + interp(t);
  }
  for (int t = threadIdx.x + PF * blockDim.x; t < nlf; t += blockDim.x)
    xf[fbase + t] = (assign ? T(0) : xf[fbase + t]) + interp(t);
}

template <typename T> int diag_launch(const T* E, const GridArgs<T>& g, T* out, int invert, hipStream_t st);
template <typename T> int energy_launch(const T* u, const GridArgs<T>& g, T* ce, const T* rho, T* dc, T penal, T dE, hipStream_t st);
template <typename T> int dense_launch(const T* u, const T* E, const T* ke, T* y, int nx, int ny, int nz, hipStream_t st);
template <typename T> int restrict_launch(const T* rf, T* rc, const GridArgs<T>& g, const uint8_t* cmask, int cmask_kind, hipStream_t st);
template <typename T> int prolong_launch(const T* ec, T* xf, const GridArgs<T>& g, int assign, hipStream_t st);
template <typename T> int coarsen_modulus_launch(const T* Ef, T* Ec, int nx, int ny, int nz, hipStream_t st);

}  // namespace vh
