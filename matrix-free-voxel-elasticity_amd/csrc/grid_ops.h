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
// (nx, ny, nz) is always the FINE element grid.  Both operators work on blocks of whole rows
// of the author layout (a row = the 3 (nz + 1) values of one (ix, iy) z-line, contiguous in HBM):
//  restriction: coarse plane Ix x TR_RB coarse rows x a z segment of coarse nodes.  The x weights
//    combine the three fine planes 2Ix - 1 .. 2Ix + 1 into 2 TR_RB + 1 weighted fine rows in LDS
//    (one coalesced pass, every load of the block in flight together); each coarse DOF then takes
//    its y and z neighbours from LDS.
//  prolongation: fine plane ix x TR_PB fine rows x a z segment of fine nodes.  The fine values
//    (for xf +=) are loaded first, then the TR_PB / 2 + 1 coarse rows are combined with the x
//    weight into LDS; each fine DOF interpolates along y and z from LDS.
// Blocks of one grid are numbered z segment fastest, then y block, then x plane.
constexpr int TR_THREADS = 256;
constexpr int TR_RB = 4;            // coarse rows per restriction block
constexpr int TR_CZ = 80;           // coarse nodes per restriction z segment (at most)
constexpr int TR_PB = 8;            // fine rows per prolongation block (even)
constexpr int TR_FZ = 160;          // fine nodes per prolongation z segment (at most; even)
constexpr int TR_PF = (TR_PB * 3 * TR_FZ + TR_THREADS - 1) / TR_THREADS;   // fine values per thread
constexpr int TR_U = 4;             // slots per thread with loads in flight together

struct TransferGeom {
  int ny1, nz1;   // nodes per column / row of the grid the blocks tile (coarse: restriction)
  int seg;        // z nodes per segment (prolongation: even, so a segment starts on a coarse node)
  int nyb, nzb;   // blocks along y, segments along z
  int nblocks;
  size_t lds;     // LDS values per block
};

__host__ __device__ inline TransferGeom restrict_geom(int nx, int ny, int nz) {
  TransferGeom g;
  g.ny1 = ny / 2 + 1;
  g.nz1 = nz / 2 + 1;
  g.nzb = (g.nz1 + TR_CZ - 1) / TR_CZ;
  g.seg = (g.nz1 + g.nzb - 1) / g.nzb;
  g.nyb = (g.ny1 + TR_RB - 1) / TR_RB;
  g.nblocks = (nx / 2 + 1) * g.nyb * g.nzb;
  g.lds = size_t(2 * TR_RB + 1) * 3 * (2 * g.seg + 1);
  return g;
}

__host__ __device__ inline TransferGeom prolong_geom(int nx, int ny, int nz) {
  TransferGeom g;
  g.ny1 = ny + 1;
  g.nz1 = nz + 1;
  g.nzb = (g.nz1 + TR_FZ - 1) / TR_FZ;
  g.seg = (((g.nz1 + g.nzb - 1) / g.nzb) + 1) & ~1;
  g.nyb = (g.ny1 + TR_PB - 1) / TR_PB;
  g.nblocks = (nx + 1) * g.nyb * g.nzb;
  g.lds = size_t(TR_PB / 2 + 1) * 3 * (g.seg / 2 + 1);
  return g;
}

// rc = mask_c * R (mask_f * rf) on restriction block `blk`; xl: restrict_geom(..).lds values.
template <typename T>
__device__ __forceinline__ void restrict_block(const T* __restrict__ rf, T* __restrict__ rc, int nx, int ny, int nz,
                                               int fmask_kind, const uint8_t* __restrict__ fmask,
                                               const uint8_t* __restrict__ cmask, int cmask_kind, int blk, T* xl) {
  const TransferGeom g = restrict_geom(nx, ny, nz);
  const int zb = blk % g.nzb;
  const int yb = (blk / g.nzb) % g.nyb;
  const int Ix = blk / (g.nzb * g.nyb);
  const int Iy0 = yb * TR_RB, Iz0 = zb * g.seg;
  const int nrow = min(TR_RB, g.ny1 - Iy0);        // coarse rows of the block
  const int ncz = min(g.seg, g.nz1 - Iz0);         // coarse nodes of the segment
  const int flen = 3 * (2 * ncz + 1);              // fine nodes 2 Iz0 - 1 .. 2 (Iz0 + ncz) - 1
  const int nslot = (2 * nrow + 1) * flen;         // fine rows 2 Iy0 - 1 .. 2 (Iy0 + nrow) - 1
  const int fz3 = 3 * (nz + 1);
  const size_t fplane = size_t(ny + 1) * fz3;
  const int q0 = 3 * (2 * Iz0 - 1);                // row offset of the segment's first value
  const int ixm = 2 * Ix - 1, ix0 = 2 * Ix, ixp = 2 * Ix + 1;
  const bool okm = ixm >= 0 && !(fmask_kind == MASK_X0 && ixm == 0);
  const bool ok0 = !(fmask_kind == MASK_X0 && ix0 == 0);
  const bool okp = ixp <= nx;
  auto fetch = [&](size_t dof) -> T {
    const T v = rf[dof];
    return (fmask_kind == MASK_ARRAY && !fmask[dof]) ? T(0) : v;
  };
  for (int s0 = threadIdx.x; s0 < nslot; s0 += TR_U * TR_THREADS) {
    T vm[TR_U], v0[TR_U], vp[TR_U];
#pragma unroll
    for (int u = 0; u < TR_U; ++u) {
      const int s = s0 + u * TR_THREADS;
      vm[u] = v0[u] = vp[u] = T(0);
      const int j = s / flen, c = s - (s / flen) * flen;
      const int iy = 2 * Iy0 - 1 + j, q = q0 + c;
      if (s < nslot && iy >= 0 && iy <= ny && q >= 0 && q < fz3) {
        const size_t off = size_t(iy) * fz3 + q;
        if (okm) vm[u] = fetch(size_t(ixm) * fplane + off);
        if (ok0) v0[u] = fetch(size_t(ix0) * fplane + off);
        if (okp) vp[u] = fetch(size_t(ixp) * fplane + off);
      }
    }
#pragma unroll
    for (int u = 0; u < TR_U; ++u) {
      const int s = s0 + u * TR_THREADS;
      if (s < nslot) xl[s] = T(0.5) * vm[u] + v0[u] + T(0.5) * vp[u];
    }
  }
  __syncthreads();
  const int clen = 3 * ncz;
  const size_t cz3 = size_t(3) * g.nz1;
  auto zsum = [&](const T* p) { return T(0.5) * p[0] + p[3] + T(0.5) * p[6]; };
  for (int o = threadIdx.x; o < nrow * clen; o += TR_THREADS) {
    const int r = o / clen, c = o - (o / clen) * clen;
    const int zl = c / 3, k = c - 3 * (c / 3);
    // coarse row r <-> local fine row 2r + 1, coarse node zl <-> local fine node 2 zl + 1
    const T* p = xl + 2 * r * flen + 6 * zl + k;
    const T acc = T(0.5) * zsum(p) + zsum(p + flen) + T(0.5) * zsum(p + 2 * flen);
    const size_t ci = (size_t(Ix) * g.ny1 + Iy0 + r) * cz3 + 3 * size_t(Iz0) + c;
    bool free_dof = true;
    if (cmask_kind == MASK_X0) free_dof = Ix > 0;
    if (cmask_kind == MASK_ARRAY) free_dof = cmask[ci] != 0;
    rc[ci] = free_dof ? acc : T(0);
  }
}

// xf (+)= mask_f * P ec on prolongation block `blk`; cl: prolong_geom(..).lds values.
template <typename T>
__device__ __forceinline__ void prolong_block(const T* __restrict__ ec, T* __restrict__ xf, int nx, int ny, int nz,
                                              int mask_kind, const uint8_t* __restrict__ mask, int assign, int blk,
                                              T* cl) {
  const TransferGeom g = prolong_geom(nx, ny, nz);
  const int zb = blk % g.nzb;
  const int yb = (blk / g.nzb) % g.nyb;
  const int ix = blk / (g.nzb * g.nyb);
  const int iy0 = yb * TR_PB, iz0 = zb * g.seg;    // both even
  const int nrow = min(TR_PB, g.ny1 - iy0);
  const int flen = 3 * min(g.seg, g.nz1 - iz0);
  const int nslot = nrow * flen;                   // <= TR_PF * TR_THREADS
  const int cy1 = ny / 2 + 1, cz1 = nz / 2 + 1;
  const int Y0 = iy0 >> 1, Z0 = iz0 >> 1;
  const int crow = min(TR_PB / 2 + 1, cy1 - Y0);
  const int clen = 3 * min(g.seg / 2 + 1, cz1 - Z0);
  const int fz3 = 3 * g.nz1;
  const size_t cz3 = size_t(3) * cz1;
  const size_t fbase = (size_t(ix) * g.ny1 + iy0) * fz3 + 3 * size_t(iz0);
  // the fine values this thread updates are loaded before the coarse rows, so both memory
  // round trips of the workgroup overlap
  T xold[TR_PF];
#pragma unroll
  for (int j = 0; j < TR_PF; ++j) {
    const int s = threadIdx.x + j * TR_THREADS;
    const int r = s / flen, c = s - (s / flen) * flen;
    xold[j] = (!assign && s < nslot) ? xf[fbase + size_t(r) * fz3 + c] : T(0);
  }
  const int x0 = ix >> 1, x1 = (ix + 1) >> 1;
  const bool xodd = ix & 1;
  const T* c0 = ec + (size_t(x0) * cy1 + Y0) * cz3 + 3 * size_t(Z0);
  const T* c1 = ec + (size_t(x1) * cy1 + Y0) * cz3 + 3 * size_t(Z0);
  const int ncs = crow * clen;
  for (int s0 = threadIdx.x; s0 < ncs; s0 += TR_U * TR_THREADS) {
    T a[TR_U], b[TR_U];
#pragma unroll
    for (int u = 0; u < TR_U; ++u) {
      const int s = s0 + u * TR_THREADS;
      const int j = s / clen, c = s - (s / clen) * clen;
      const size_t off = size_t(j) * cz3 + c;
      a[u] = s < ncs ? c0[off] : T(0);
      b[u] = (s < ncs && xodd) ? c1[off] : T(0);
    }
#pragma unroll
    for (int u = 0; u < TR_U; ++u) {
      const int s = s0 + u * TR_THREADS;
      if (s < ncs) cl[s] = xodd ? T(0.5) * (a[u] + b[u]) : a[u];
    }
  }
  __syncthreads();
  const bool plane_free = !(mask_kind == MASK_X0 && ix == 0);
#pragma unroll
  for (int j = 0; j < TR_PF; ++j) {
    const int s = threadIdx.x + j * TR_THREADS;
    if (s < nslot) {
      const int r = s / flen, c = s - (s / flen) * flen;
      const int zl = c / 3, k = c - 3 * (c / 3);
      const T* p = cl + (r >> 1) * clen + 3 * (zl >> 1) + k;
      T v = p[0];
      if (zl & 1) v = T(0.5) * (v + p[3]);
      if (r & 1) {
        T w = p[clen];
        if (zl & 1) w = T(0.5) * (w + p[clen + 3]);
        v = T(0.5) * (v + w);
      }
      const size_t fi = fbase + size_t(r) * fz3 + c;
      bool free_dof = plane_free;
      if (mask_kind == MASK_ARRAY) free_dof = mask[fi] != 0;
      xf[fi] = (assign ? T(0) : xold[j]) + (free_dof ? v : T(0));
    }
  }
}

template <typename T> int diag_launch(const T* E, const GridArgs<T>& g, T* out, int invert, hipStream_t st);
template <typename T> int energy_launch(const T* u, const GridArgs<T>& g, T* ce, const T* rho, T* dc, T penal, T dE, hipStream_t st);
template <typename T> int dense_launch(const T* u, const T* E, const T* ke, T* y, int nx, int ny, int nz, hipStream_t st);
template <typename T> int restrict_launch(const T* rf, T* rc, const GridArgs<T>& g, const uint8_t* cmask, int cmask_kind, hipStream_t st);
template <typename T> int prolong_launch(const T* ec, T* xf, const GridArgs<T>& g, int assign, hipStream_t st);
int assemble_dense_launch(const double* ke, const double* E, int nx, int ny, int nz, int mask_kind,
                          const uint8_t* mask, double* A, int ld, hipStream_t st);
int galerkin_scatter_launch(const double* G, int nx, int ny, int nz, int mask_kind, const uint8_t* mask, double* A,
                            int ld, hipStream_t st);
template <typename T> int coarsen_modulus_launch(const T* Ef, T* Ec, int nx, int ny, int nz, hipStream_t st);

}  // namespace vh
