// Matrix-free Hex8 voxel elasticity product y = K(E) u and its fused epilogues.
//
// Algebra (code/hex_modal/modal.py:9-16,55-77 of the reference): per element, with the
// 8 nodes in binary order l = x + 2y + 4z, the modal coefficients v_p = sum_l H[p,l] u_l
// (H[p,l] = (-1)^popcount(p&l), unnormalised) turn the 24x24 element matrix into 8 blocks
// of 3x3 (C = T K T^T / 64): block g couples (mode g^1, x), (mode g^2, y), (mode g^4, z),
// and the nodal forces are f_l = sum_p H[l,p] (E_e C v)_p.
//
// MI355X mapping (plane sweep, no atomics, deterministic):
//  * A workgroup owns a TY x TZ tile of node columns (y, z) and a chunk of LX node planes
//    in x, the slowest author axis.  It marches element planes ex = x0-1 .. x1-1 in x.
//  * Each of its (TY+1) x (TZ+1) threads owns one element column (ey, ez): the halo row and
//    column of elements on the low side are recomputed so every owned node receives all
//    eight element contributions inside the workgroup and y is written exactly once.
//  * The 8-point transform is factored: the 2x2 (y,z) transform of a node face is computed
//    once per plane and carried in registers to the next element along x (each face is
//    shared by elements ex-1 and ex); the x stage is a sum/difference of two faces.  The
//    inverse mirrors it: the left-face half of element ex is added to the carried right-face
//    half of element ex-1 before the (y,z) inverse.  Per element this is 48 + 72 + 24 + 48
//    + 12 FP64 operations (forward, blocks, modulus, inverse, carry) instead of 240 with an
//    8-lane shuffle butterfly.
//  * Node planes are staged through LDS (SoA, coalesced flat loads with a one-plane
//    register prefetch); the four corner contributions of each node are gathered from LDS in
//    a fixed order, so results are bitwise reproducible.
//  * The epilogue fuses the constrained operator (M K M + I - M), the residual, a damped
//    Jacobi sweep, a Chebyshev step or a p.Ap partial sum into the single pass over y.
//  * Workgroup ids are remapped so that each XCD (8 per MI355X, own L2) receives a contiguous
//    range of tiles with the x chunk slowest: halo rows, columns and planes shared by
//    neighbouring tiles are then mostly read from the same L2.
#include "common.h"
#include "matvec.h"

namespace vh {

template <typename T, typename S, int MODE, int MASK>
__global__ void __launch_bounds__(MV_BLOCK) parity_xmarch(MatvecParams<T> a) {
  __shared__ T su[3][MV_PLANE];          // staged node plane (masked input), SoA
  __shared__ T sc[4][3][MV_PLANE];       // corner contributions, indexed by node
  __shared__ double red[MV_BLOCK / 64];

  // ---- XCD-aware tile decode: hardware assigns block b to XCD b % 8; give XCD k the
  // contiguous logical range [k*per, (k+1)*per).
  const int per = gridDim.x >> 3;
  const int b = blockIdx.x;
  const int logical = (b & 7) * per + (b >> 3);
  if (logical >= a.n_tiles) return;
  const int tz = logical % a.tiles_z;
  const int ty = (logical / a.tiles_z) % a.tiles_y;
  const int cx = logical / (a.tiles_z * a.tiles_y);

  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int ny1 = ny + 1, nz1 = nz + 1;
  const int oy0 = ty * MV_TY, oy1 = (ty == a.tiles_y - 1) ? ny1 : oy0 + MV_TY;
  const int oz0 = tz * MV_TZ, oz1 = (tz == a.tiles_z - 1) ? nz1 : oz0 + MV_TZ;
  const int ox0 = cx * a.lx, ox1 = (cx == a.chunks_x - 1) ? nx + 1 : ox0 + a.lx;
  const int ybase = oy0 - 1, zbase = oz0 - 1;      // local node j <-> global base + j

  const int tid = threadIdx.x;
  const int ly = tid / MV_EZ, lz = tid - (tid / MV_EZ) * MV_EZ;
  const int ey = ybase + ly, ez = zbase + lz;
  const bool col_in_tile = tid < MV_NCOL;
  const bool col_ok = col_in_tile && ey >= 0 && ey < ny && ez >= 0 && ez < nz;
  const int iy_own = oy0 + ly, iz_own = oz0 + lz;  // node (ly+1, lz+1) in local coordinates
  const bool node_own = col_in_tile && iy_own < oy1 && iz_own < oz1;

  // zero the contribution slots once: slots without a writer column must read as zero
  {
    T* flat = &sc[0][0][0];
    for (int i = tid; i < 12 * MV_PLANE; i += MV_BLOCK) flat[i] = T(0);
  }

  const S* __restrict__ u = reinterpret_cast<const S*>(a.u);

  // ---- flat, coalesced plane loader: NY rows of 3*NZ contiguous values per plane.
  T pre[MV_LOADS];
  auto prefetch = [&](int ix) {
#pragma unroll
    for (int j = 0; j < MV_LOADS; ++j) {
      const int k = tid + j * MV_BLOCK;
      T v = T(0);
      if (k < MV_NY * 3 * MV_NZ) {
        const int row = k / (3 * MV_NZ);
        const int col = k - row * (3 * MV_NZ);
        const int iy = ybase + row;
        const int iz = zbase + col / 3;
        if (ix >= 0 && ix <= nx && iy >= 0 && iy <= ny && iz >= 0 && iz <= nz) {
          const size_t dof = 3 * ((size_t(ix) * ny1 + iy) * nz1 + zbase) + col;
          bool free_dof = true;
          if (MASK == MASK_X0) free_dof = ix > 0;
          if (MASK == MASK_ARRAY) free_dof = a.mask[dof] != 0;
          if (free_dof) v = load_as<S, T>(u, dof);
        }
      }
      pre[j] = v;
    }
  };
  auto commit = [&]() {
#pragma unroll
    for (int j = 0; j < MV_LOADS; ++j) {
      const int k = tid + j * MV_BLOCK;
      if (k < MV_NY * 3 * MV_NZ) {
        const int row = k / (3 * MV_NZ);
        const int col = k - row * (3 * MV_NZ);
        const int jz = col / 3, c = col - 3 * (col / 3);
        su[c][row * MV_NZ + jz] = pre[j];
      }
    }
  };
  // 2x2 (y,z) transform of this column's face: F[q], q = py + 2 pz.
  auto face = [&](T (&F)[4][3]) {
    const int i00 = ly * MV_NZ + lz;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T u00 = su[c][i00], u01 = su[c][i00 + 1];
      const T u10 = su[c][i00 + MV_NZ], u11 = su[c][i00 + MV_NZ + 1];
      const T s0 = u00 + u01, d0 = u00 - u01, s1 = u10 + u11, d1 = u10 - u11;
      F[0][c] = s0 + s1;
      F[1][c] = s0 - s1;
      F[2][c] = d0 + d1;
      F[3][c] = d0 - d1;
    }
  };

  T Fl[4][3], Fr[4][3], Gc[4][3];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int c = 0; c < 3; ++c) Gc[q][c] = T(0);

  double dot_acc = 0.0;
  const T* __restrict__ E = a.E;

  prefetch(ox0 - 1);
  __syncthreads();                 // sc zeroed
  commit();
  __syncthreads();
  prefetch(ox0);
  face(Fl);

  for (int ex = ox0 - 1; ex < ox1; ++ex) {
    __syncthreads();               // all face() reads of the previous plane are done
    commit();                      // node plane ex + 1
    __syncthreads();
    prefetch(ex + 2);
    face(Fr);

    T Gl[4][3], Gr[4][3];
    if (col_ok && ex >= 0 && ex < nx) {
      const T Ee = E[(size_t(ex) * ny + ey) * nz + ez];
      T v[8][3];
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          v[2 * q][c] = Fl[q][c] + Fr[q][c];
          v[2 * q + 1][c] = Fl[q][c] - Fr[q][c];
        }
      T f[8][3];
#pragma unroll
      for (int g = 0; g < 8; ++g) {
        const T a0 = Ee * v[g ^ 1][0], a1 = Ee * v[g ^ 2][1], a2 = Ee * v[g ^ 4][2];
        f[g ^ 1][0] = a.coef[9 * g + 0] * a0 + a.coef[9 * g + 1] * a1 + a.coef[9 * g + 2] * a2;
        f[g ^ 2][1] = a.coef[9 * g + 3] * a0 + a.coef[9 * g + 4] * a1 + a.coef[9 * g + 5] * a2;
        f[g ^ 4][2] = a.coef[9 * g + 6] * a0 + a.coef[9 * g + 7] * a1 + a.coef[9 * g + 8] * a2;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          Gl[q][c] = f[2 * q][c] + f[2 * q + 1][c];
          Gr[q][c] = f[2 * q][c] - f[2 * q + 1][c];
        }
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) Gl[q][c] = Gr[q][c] = T(0);
    }

    if (ex >= ox0) {
      // node plane ex is complete: carried right half of element ex-1 + left half of ex
      if (col_in_tile) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const T g0 = Gc[0][c] + Gl[0][c], g1 = Gc[1][c] + Gl[1][c];
          const T g2 = Gc[2][c] + Gl[2][c], g3 = Gc[3][c] + Gl[3][c];
          const T A0 = g0 + g1, A1 = g0 - g1, B0 = g2 + g3, B1 = g2 - g3;
          const int i00 = ly * MV_NZ + lz;
          sc[0][c][i00] = A0 + B0;                       // corner (b=0, c=0)
          sc[1][c][i00 + 1] = A0 - B0;                   // corner (0, 1)
          sc[2][c][i00 + MV_NZ] = A1 + B1;               // corner (1, 0)
          sc[3][c][i00 + MV_NZ + 1] = A1 - B1;           // corner (1, 1)
        }
      }
      __syncthreads();
      if (node_own) {
        const int jn = (ly + 1) * MV_NZ + (lz + 1);
        const size_t node = (size_t(ex) * ny1 + iy_own) * nz1 + iz_own;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const T ku = ((sc[0][c][jn] + sc[1][c][jn]) + sc[2][c][jn]) + sc[3][c][jn];
          const size_t dof = 3 * node + c;
          bool free_dof = true;
          if (MASK == MASK_X0) free_dof = ex > 0;
          if (MASK == MASK_ARRAY) free_dof = a.mask[dof] != 0;
          // unmasked input at the owned DOF (the staged plane already moved on to ex + 1);
          // the re-read hits the cache, the plane was fetched two steps earlier
          T uo = T(0);
          if (MODE != MODE_RAW || MASK != MASK_NONE) uo = load_as<S, T>(u, dof);
          // constrained operator (M K M + I - M) u, or the raw K u without a mask
          const T au = (MASK == MASK_NONE || free_dof) ? ku : (a.fixed_zero ? T(0) : uo);
          T out;
          if (MODE == MODE_RAW) {
            out = au;
          } else if (MODE == MODE_RESID) {
            out = a.b[dof] - au;
          } else if (MODE == MODE_JACOBI) {
            out = uo + a.omega * a.dinv[dof] * (a.b[dof] - au);
          } else if (MODE == MODE_CHEB) {
            const T z = a.dinv[dof] * (a.b[dof] - au);
            const T d = a.cheb_a * (a.cheb_c == T(0) ? z : z + a.cheb_c * a.d[dof]);
            a.d[dof] = d;
            out = uo + d;
          } else {  // MODE_DOT: y = A u and sum u . A u
            out = au;
            dot_acc += double(uo) * double(au);
          }
          if (MASK != MASK_NONE && a.fixed_zero && !free_dof) out = T(0);
          a.y[dof] = out;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Gc[q][c] = Gr[q][c];
        Fl[q][c] = Fr[q][c];
      }
  }

  if (MODE == MODE_DOT) {
    const double s = block_sum<double, MV_BLOCK>(dot_acc, red);
    if (tid == 0) a.partial[logical] = s;
  }
}

// ------------------------------------------------------------------------------ launcher
template <typename T, typename S, int MODE, int MASK>
static void launch_t(const MatvecParams<T>& p, int grid, hipStream_t st) {
  hipLaunchKernelGGL((parity_xmarch<T, S, MODE, MASK>), dim3(grid), dim3(MV_BLOCK), 0, st, p);
}

template <typename T, typename S, int MODE>
static void launch_mask(const MatvecParams<T>& p, int mask_kind, int grid, hipStream_t st) {
  switch (mask_kind) {
    case MASK_NONE: launch_t<T, S, MODE, MASK_NONE>(p, grid, st); break;
    case MASK_ARRAY: launch_t<T, S, MODE, MASK_ARRAY>(p, grid, st); break;
    default: launch_t<T, S, MODE, MASK_X0>(p, grid, st); break;
  }
}

template <typename T, typename S>
static void launch_mode(const MatvecParams<T>& p, int mode, int mask_kind, int grid, hipStream_t st) {
  switch (mode) {
    case MODE_RAW: launch_mask<T, S, MODE_RAW>(p, mask_kind, grid, st); break;
    case MODE_RESID: launch_mask<T, S, MODE_RESID>(p, mask_kind, grid, st); break;
    case MODE_JACOBI: launch_mask<T, S, MODE_JACOBI>(p, mask_kind, grid, st); break;
    case MODE_CHEB: launch_mask<T, S, MODE_CHEB>(p, mask_kind, grid, st); break;
    default: launch_mask<T, S, MODE_DOT>(p, mask_kind, grid, st); break;
  }
}

void matvec_geometry(int nx, int ny, int nz, int lx_hint, int* tiles_y, int* tiles_z,
                     int* chunks_x, int* lx, int* n_tiles, int* grid) {
  const int ty = (ny + MV_TY - 1) / MV_TY;
  const int tz = (nz + MV_TZ - 1) / MV_TZ;
  int l = lx_hint;
  if (l <= 0) {
    // aim for >= 2 resident workgroups per CU on 256 CUs, with chunks of at least 8 planes
    const int want = 1024;
    const int per_plane = ty * tz;
    int chunks = (want + per_plane - 1) / per_plane;
    l = (nx + chunks - 1) / chunks;
    if (l < 8) l = 8;
  }
  if (l > nx) l = nx;
  if (l < 1) l = 1;
  const int cx = (nx + l - 1) / l;
  *tiles_y = ty;
  *tiles_z = tz;
  *chunks_x = cx;
  *lx = l;
  *n_tiles = ty * tz * cx;
  *grid = ((*n_tiles + 7) / 8) * 8;
}

template <typename T>
int matvec_launch(MatvecParams<T> p, int storage, int mode, int mask_kind, hipStream_t st) {
  int grid = 0;
  matvec_geometry(p.nx, p.ny, p.nz, p.lx, &p.tiles_y, &p.tiles_z, &p.chunks_x, &p.lx,
                  &p.n_tiles, &grid);
  if (storage == BF16) {
    if constexpr (sizeof(T) == 4) launch_mode<T, uint16_t>(p, mode, mask_kind, grid, st);
    else return fail(-2, "bf16 storage requires fp32 compute");
  } else {
    launch_mode<T, T>(p, mode, mask_kind, grid, st);
  }
  return hip_status(hipGetLastError(), "parity_xmarch launch");
}

template int matvec_launch<double>(MatvecParams<double>, int, int, int, hipStream_t);
template int matvec_launch<float>(MatvecParams<float>, int, int, int, hipStream_t);

}  // namespace vh
