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
//  restriction: coarse plane Ix x R coarse rows x a z segment of coarse nodes.  The x weights
//    combine the three fine planes 2Ix - 1 .. 2Ix + 1 into 2 R + 1 weighted fine rows in LDS
//    (one coalesced pass, every load of the block in flight together); each coarse DOF then takes
//    its y and z neighbours from LDS.
//  prolongation: fine plane ix x R fine rows x a z segment of fine nodes.  The fine values
//    (for xf +=) are loaded first, then the R / 2 + 1 coarse rows are combined with the x
//    weight into LDS; each fine DOF interpolates along y and z from LDS.
// Blocks of one grid are numbered z segment fastest, then y block, then x plane.  The workgroup
// size and the largest z segment are compile-time configurations (RestrictCfg / ProlongCfg; the
// library's kernels use TR_RESTRICT and TR_PROLONG); the rows per block R are RB / PB, or more on
// grids whose z segments are short (restrict_rows, prolong_rows).
template <int RB_, int CZ_, int NT_, int U_>
struct RestrictCfg {
  static constexpr int RB = RB_;   // coarse rows per block
  static constexpr int CZ = CZ_;   // coarse nodes per z segment (at most)
  static constexpr int NT = NT_;   // threads per workgroup
  static constexpr int SLOTS = ((2 * RB + 1) * (2 * CZ + 1) + NT - 1) / NT;     // fine nodes per thread
  static constexpr int U = U_ < SLOTS ? U_ : SLOTS;                             // nodes with loads in flight
};
template <int PB_, int FZ_, int NT_>
struct ProlongCfg {
  static constexpr int PB = PB_;   // fine rows per block (even)
  static constexpr int FZ = FZ_;   // fine nodes per z segment (at most; even)
  static constexpr int NT = NT_;
  static constexpr int PF = (PB * FZ + NT - 1) / NT;                              // fine nodes per thread
  static constexpr int CS = ((PB / 2 + 1) * (FZ / 2 + 1) + NT - 1) / NT;          // coarse nodes per thread
};
using TR_RESTRICT = RestrictCfg<4, 80, 256, 6>;
using TR_PROLONG = ProlongCfg<4, 160, 256>;

// Walks the slots s = start, start + step, ... of a block laid out as rows of `len` values,
// tracking (row, col) = (s / len, s % len) without a division per slot.
struct RowCursor {
  int r, c, dr, dc, len;
  __device__ RowCursor(int start, int step, int len_) : len(len_) {
    r = start / len;
    c = start - r * len;
    dr = step / len;
    dc = step - dr * len;
  }
  __device__ void next() {
    r += dr;
    c += dc;
    if (c >= len) { c -= len; ++r; }
  }
};

struct TransferGeom {
  int ny1, nz1;   // nodes per column / row of the grid the blocks tile (coarse: restriction)
  int seg;        // z nodes per segment (prolongation: even, so a segment starts on a coarse node)
  int rows;       // rows per block (restriction: coarse rows; prolongation: fine rows, even)
  int nyb, nzb;   // blocks along y, segments along z
  int nblocks;
  size_t lds;     // LDS values per block
};

template <typename C>
__host__ __device__ inline TransferGeom restrict_geom(int nx, int ny, int nz, int rows) {
  TransferGeom g;
  g.ny1 = ny / 2 + 1;
  g.nz1 = nz / 2 + 1;
  g.nzb = (g.nz1 + C::CZ - 1) / C::CZ;
  g.seg = (g.nz1 + g.nzb - 1) / g.nzb;
  g.rows = rows;
  g.nyb = (g.ny1 + g.rows - 1) / g.rows;
  g.nblocks = (nx / 2 + 1) * g.nyb * g.nzb;
  g.lds = size_t(2 * g.rows + 1) * 3 * (2 * g.seg + 1);
  return g;
}

template <typename C>
__host__ __device__ inline TransferGeom prolong_geom(int nx, int ny, int nz, int rows) {
  TransferGeom g;
  g.ny1 = ny + 1;
  g.nz1 = nz + 1;
  g.nzb = (g.nz1 + C::FZ - 1) / C::FZ;
  g.seg = (((g.nz1 + g.nzb - 1) / g.nzb) + 1) & ~1;
  g.rows = rows;
  g.nyb = (g.ny1 + g.rows - 1) / g.rows;
  g.nblocks = (nx + 1) * g.nyb * g.nzb;
  g.lds = size_t(g.rows / 2 + 1) * 3 * (g.seg / 2 + 1);
  return g;
}

// Rows per block (host side; the kernels take the result as an argument).  RB / PB, or more
// where the z segment is short (narrow grids, coarse levels, the small problems of a batch), so
// that a block still stages about U NT fine nodes (restriction: one round of loads) or fills its
// PF fine and CS coarse nodes per thread (prolongation: rows seg <= PF NT and (rows / 2 + 1)
// (seg / 2 + 1) <= CS NT; PB meets both since seg <= FZ).  The rows grow only while the launch
// keeps at least min_blocks blocks (0: no limit, for a batch whose problems fill the chip
// together).
template <typename C>
inline int restrict_rows(int nx, int ny, int nz, int min_blocks) {
  const TransferGeom g = restrict_geom<C>(nx, ny, nz, C::RB);
  const int fnodes = 2 * g.seg + 1;
  int rows = ((C::U * C::NT) / fnodes - 1) / 2;
  while (rows > C::RB && restrict_geom<C>(nx, ny, nz, rows).nblocks < min_blocks) --rows;
  return rows > C::RB ? rows : C::RB;
}

template <typename C>
inline int prolong_rows(int nx, int ny, int nz, int min_blocks) {
  const TransferGeom g = prolong_geom<C>(nx, ny, nz, C::PB);
  const int by_fine = (C::PF * C::NT) / g.seg;
  const int by_coarse = 2 * ((C::CS * C::NT) / (g.seg / 2 + 1)) - 2;
  int rows = (by_fine < by_coarse ? by_fine : by_coarse) & ~1;
  while (rows > C::PB && prolong_geom<C>(nx, ny, nz, rows).nblocks < min_blocks) rows -= 2;
  return rows > C::PB ? rows : C::PB;
}

// rc = mask_c * R (mask_f * rf) on restriction block `blk` of the geometry with `rows` coarse rows
// per block; xl: restrict_geom(..).lds values.
template <typename T, typename C = TR_RESTRICT>
__device__ __forceinline__ void restrict_block(const T* __restrict__ rf, T* __restrict__ rc, int nx, int ny, int nz,
                                               int fmask_kind, const uint8_t* __restrict__ fmask,
                                               const uint8_t* __restrict__ cmask, int cmask_kind, int rows, int blk,
                                               T* xl) {
  constexpr int NT = C::NT, U = C::U;
  const TransferGeom g = restrict_geom<C>(nx, ny, nz, rows);
  const int zb = blk % g.nzb;
  const int yb = (blk / g.nzb) % g.nyb;
  const int Ix = blk / (g.nzb * g.nyb);
  const int Iy0 = yb * g.rows, Iz0 = zb * g.seg;
  const int nrow = min(g.rows, g.ny1 - Iy0);       // coarse rows of the block
  const int ncz = min(g.seg, g.nz1 - Iz0);         // coarse nodes of the segment
  const int fnodes = 2 * ncz + 1;                  // fine nodes 2 Iz0 - 1 .. 2 (Iz0 + ncz) - 1
  const int flen = 3 * fnodes;
  const int nslot = (2 * nrow + 1) * fnodes;       // fine rows 2 Iy0 - 1 .. 2 (Iy0 + nrow) - 1
  const int fz1 = nz + 1;
  const size_t fplane = size_t(ny + 1) * 3 * fz1;
  const int n0 = 2 * Iz0 - 1;                      // the segment's first fine node
  const int ixm = 2 * Ix - 1, ix0 = 2 * Ix, ixp = 2 * Ix + 1;
  const bool okm = ixm >= 0 && !(fmask_kind == MASK_X0 && ixm == 0);
  const bool ok0 = !(fmask_kind == MASK_X0 && ix0 == 0);
  const bool okp = ixp <= nx;
  // a slot is one fine node of the block's rows: its three values are read by adjacent loads
  // (one wide load per node and plane), every load of a round in flight together
  RowCursor cur(threadIdx.x, NT, fnodes);
  for (int s0 = threadIdx.x; s0 < nslot; s0 += U * NT) {
    T vm[U][3], v0[U][3], vp[U][3];
    size_t off[U];
    bool in[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u * NT;
      const int iy = 2 * Iy0 - 1 + cur.r, iz = n0 + cur.c;
      cur.next();
      in[u] = s < nslot && iy >= 0 && iy <= ny && iz >= 0 && iz <= nz;
      off[u] = in[u] ? (size_t(iy) * fz1 + iz) * 3 : 0;
#pragma unroll
      for (int c = 0; c < 3; ++c) vm[u][c] = v0[u][c] = vp[u][c] = T(0);
      if (in[u]) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          if (okm) vm[u][c] = rf[size_t(ixm) * fplane + off[u] + c];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          if (ok0) v0[u][c] = rf[size_t(ix0) * fplane + off[u] + c];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          if (okp) vp[u][c] = rf[size_t(ixp) * fplane + off[u] + c];
        }
      }
    }
    if (fmask_kind == MASK_ARRAY) {                // the fine mask, applied to the loaded values
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (in[u]) {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            if (okm && !fmask[size_t(ixm) * fplane + off[u] + c]) vm[u][c] = T(0);
            if (ok0 && !fmask[size_t(ix0) * fplane + off[u] + c]) v0[u][c] = T(0);
            if (okp && !fmask[size_t(ixp) * fplane + off[u] + c]) vp[u][c] = T(0);
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u * NT;
      if (s < nslot) {
#pragma unroll
        for (int c = 0; c < 3; ++c) xl[3 * s + c] = T(0.5) * vm[u][c] + v0[u][c] + T(0.5) * vp[u][c];
      }
    }
  }
  __syncthreads();
  const int clen = 3 * ncz;
  const size_t cz3 = size_t(3) * g.nz1;
  auto zsum = [&](const T* p) { return T(0.5) * p[0] + p[3] + T(0.5) * p[6]; };
  for (RowCursor oc(threadIdx.x, NT, clen); oc.r < nrow; oc.next()) {
    const int r = oc.r, c = oc.c;
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

// xf (+)= mask_f * P ec on prolongation block `blk` of the geometry with `rows` fine rows per block
// (even); cl: prolong_geom(..).lds values.
template <typename T, typename C = TR_PROLONG>
__device__ __forceinline__ void prolong_block(const T* __restrict__ ec, T* __restrict__ xf, int nx, int ny, int nz,
                                              int mask_kind, const uint8_t* __restrict__ mask, int assign, int rows,
                                              int blk, T* cl) {
  constexpr int NT = C::NT, PF = C::PF, CS = C::CS;
  const TransferGeom g = prolong_geom<C>(nx, ny, nz, rows);
  const int zb = blk % g.nzb;
  const int yb = (blk / g.nzb) % g.nyb;
  const int ix = blk / (g.nzb * g.nyb);
  const int iy0 = yb * g.rows, iz0 = zb * g.seg;  // both even
  const int nrow = min(g.rows, g.ny1 - iy0);
  const int fn = min(g.seg, g.nz1 - iz0);          // fine nodes of the segment
  const int cy1 = ny / 2 + 1, cz1 = nz / 2 + 1;
  const int Y0 = iy0 >> 1, Z0 = iz0 >> 1;
  const int crow = min(g.rows / 2 + 1, cy1 - Y0);
  const int cn = min(g.seg / 2 + 1, cz1 - Z0);     // coarse nodes of a combined row
  const int clen = 3 * cn;
  const int fz3 = 3 * g.nz1;
  const size_t cz3 = size_t(3) * cz1;
  const size_t fbase = (size_t(ix) * g.ny1 + iy0) * fz3 + 3 * size_t(iz0);
  // every load of the block is issued before the first wait: the fine nodes this thread
  // updates, then the coarse nodes of the x-combined rows; a slot is one node, its three
  // values read (and the fine ones written) by adjacent accesses, one wide access per node
  T xold[PF][3];
  int fo[PF];                                      // offsets of this thread's fine nodes
  {
    RowCursor fc(threadIdx.x, NT, fn);
#pragma unroll
    for (int j = 0; j < PF; ++j) {
      fo[j] = fc.r < nrow ? fc.r * fz3 + 3 * fc.c : -1;
      fc.next();
    }
  }
#pragma unroll
  for (int j = 0; j < PF; ++j) {
    const bool ld = !assign && fo[j] >= 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) xold[j][c] = T(0);
    if (ld) {
#pragma unroll
      for (int c = 0; c < 3; ++c) xold[j][c] = xf[fbase + fo[j] + c];
    }
  }
  const int x0 = ix >> 1, x1 = (ix + 1) >> 1;
  const bool xodd = ix & 1;
  const T* c0 = ec + (size_t(x0) * cy1 + Y0) * cz3 + 3 * size_t(Z0);
  const T* c1 = ec + (size_t(x1) * cy1 + Y0) * cz3 + 3 * size_t(Z0);
  const int ncs = crow * cn;                       // <= CS * NT
  T ca[CS][3], cb[CS][3];
  {
    RowCursor cc(threadIdx.x, NT, cn);
#pragma unroll
    for (int u = 0; u < CS; ++u) {
      const int s = threadIdx.x + u * NT;
      const size_t off = size_t(cc.r) * cz3 + 3 * cc.c;
      cc.next();
#pragma unroll
      for (int c = 0; c < 3; ++c) ca[u][c] = cb[u][c] = T(0);
      if (s < ncs) {
#pragma unroll
        for (int c = 0; c < 3; ++c) ca[u][c] = c0[off + c];
      }
      if (s < ncs && xodd) {
#pragma unroll
        for (int c = 0; c < 3; ++c) cb[u][c] = c1[off + c];
      }
    }
  }
#pragma unroll
  for (int u = 0; u < CS; ++u) {
    const int s = threadIdx.x + u * NT;
    if (s < ncs) {
#pragma unroll
      for (int c = 0; c < 3; ++c) cl[3 * s + c] = xodd ? T(0.5) * (ca[u][c] + cb[u][c]) : ca[u][c];
    }
  }
  __syncthreads();
  const bool plane_free = !(mask_kind == MASK_X0 && ix == 0);
  RowCursor fc(threadIdx.x, NT, fn);
#pragma unroll
  for (int j = 0; j < PF; ++j) {
    const int r = fc.r, zl = fc.c;
    fc.next();
    if (fo[j] >= 0) {
      T out[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const T* p = cl + (r >> 1) * clen + 3 * (zl >> 1) + k;
        T v = p[0];
        if (zl & 1) v = T(0.5) * (v + p[3]);
        if (r & 1) {
          T w = p[clen];
          if (zl & 1) w = T(0.5) * (w + p[clen + 3]);
          v = T(0.5) * (v + w);
        }
        const size_t fi = fbase + fo[j] + k;
        bool free_dof = plane_free;
        if (mask_kind == MASK_ARRAY) free_dof = mask[fi] != 0;
        out[k] = (assign ? T(0) : xold[j][k]) + (free_dof ? v : T(0));
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) xf[fbase + fo[j] + k] = out[k];
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
