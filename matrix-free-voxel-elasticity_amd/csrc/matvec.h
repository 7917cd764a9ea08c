// Parameters and launch geometry of the plane-sweep parity matvec.
#pragma once
#include "common.h"

namespace vh {

// Owned node tile per workgroup in (y, z); element columns computed = (TY+1) x (TZ+1).
constexpr int MV_TY = 16;
constexpr int MV_TZ = 32;
constexpr int MV_EY = MV_TY + 1;
constexpr int MV_EZ = MV_TZ + 1;
constexpr int MV_NY = MV_TY + 2;                  // node rows staged per plane
constexpr int MV_NZ = MV_TZ + 2;                  // node columns staged per plane
constexpr int MV_NCOL = MV_EY * MV_EZ;            // 561 element columns
constexpr int MV_BLOCK = ((MV_NCOL + 63) / 64) * 64;   // 576 threads = 9 waves
constexpr int MV_PLANE = MV_NY * MV_NZ;           // 612 nodes
constexpr int MV_LOADS = (MV_NY * 3 * MV_NZ + MV_BLOCK - 1) / MV_BLOCK;

// Epilogue modes (A = constrained operator M K M + I - M when a mask is given, else K).
enum Mode : int {
  MODE_RAW = 0,     // y = A u
  MODE_RESID = 1,   // y = b - A u
  MODE_JACOBI = 2,  // y = u + omega * dinv * (b - A u)
  MODE_CHEB = 3,    // d = cheb_a * (dinv (b - A u) + cheb_c * d);  y = u + d
  MODE_DOT = 4,     // y = A u;  partial[tile] = sum over owned DOFs of u . A u
};

template <typename T>
struct MatvecParams {
  const void* u;          // input, storage type (T, or bf16 bits with T = float)
  const T* E;             // element moduli, author order
  T* y;                   // output (never aliases u)
  const T* b;             // right-hand side (RESID, JACOBI, CHEB)
  const T* dinv;          // inverse diagonal (JACOBI, CHEB)
  T* d;                   // Chebyshev direction, updated in place (CHEB)
  const uint8_t* mask;    // per-DOF free flags (MASK_ARRAY)
  double* partial;        // per-tile FP64 partial sums (DOT), length >= n_tiles
  int nx, ny, nz;
  int lx;                 // x chunk length (0: automatic)
  int fixed_zero;         // with a mask: the output is 0 at fixed DOFs (M K M u, (b - A u) M)
  int tiles_y, tiles_z, chunks_x, n_tiles;
  T omega, cheb_a, cheb_c;
  T coef[72];             // parity blocks C[g][i][j]
};

void matvec_geometry(int nx, int ny, int nz, int lx_hint, int* tiles_y, int* tiles_z,
                     int* chunks_x, int* lx, int* n_tiles, int* grid);

template <typename T>
int matvec_launch(MatvecParams<T> p, int storage, int mode, int mask_kind, hipStream_t st);

}  // namespace vh
