// Parameters and launch geometry of the plane-sweep parity matvec.
#pragma once
#include "common.h"

namespace vh {

// Owned node tile per workgroup in (y, z); element columns computed = (TY+1) x (TZ+1).
template <int TY_, int TZ_>
struct Tile {
  static constexpr int TY = TY_, TZ = TZ_;
  static constexpr int EY = TY + 1, EZ = TZ + 1;          // element columns (with the low halo)
  static constexpr int NY = TY + 2, NZ = TZ + 2;          // staged node rows / columns per plane
  static constexpr int NCOL = EY * EZ;
  static constexpr int BLOCK = ((NCOL + 63) / 64) * 64;
  static constexpr int PLANE = NY * NZ;
};
constexpr int MV_PREV_ZERO = 1;
constexpr int MV_W_F32 = 2;       // the AXPY operand w of an FP64 pass is stored in FP32
// load-time operations on the input (template parameter LOADOP of the product kernel); AXPY32 is
// AXPY with an FP32 w in an FP64 pass (the RESID pass of the mixed-precision Krylov loop, whose
// search direction is formed in FP32)
constexpr int LOAD_NONE = 0, LOAD_AXPY = 1, LOAD_MUL = 2, LOAD_AXPY32 = 3;
// tile variants (owned nodes in y x z; the y extent is an upper bound, the plan spreads ny evenly
// over the y tiles): 0 = automatic, 1 = 16 x 32 (576 threads), 2 = 8 x 32 (320 threads), 3 = 4 x 8
// (64 threads: one wave per workgroup, for the latency-bound coarse levels), 4 = the wide tile that
// fills one CU with a single workgroup spread evenly over its four SIMDs: 22 x 32 in FP64 (768
// threads, three waves per SIMD), 30 x 32 in FP32 (1024 threads, four waves per SIMD); 5 = 16 x 16
// (320 threads; batched products only: it wastes fewer lanes on the narrow z extents of many
// small problems)
constexpr int MV_TILE_AUTO = 0, MV_TILE_16x32 = 1, MV_TILE_8x32 = 2, MV_TILE_4x8 = 3, MV_TILE_WIDE = 4,
              MV_TILE_16x16 = 5;

// Epilogue modes (A = constrained operator M K M + I - M when a mask is given, else K).
enum Mode : int {
  MODE_RAW = 0,     // y = A u
  MODE_RESID = 1,   // y = b - A u
  MODE_JACOBI = 2,  // y = u + omega * dinv * (b - A u)
  MODE_CHEB = 3,    // d = cheb_a * (dinv (b - A u) + cheb_c * d);  y = u + d
  MODE_DOT = 4,     // y = A u;  partial[tile] = sum over owned DOFs of u . A u
  MODE_CHEB3 = 5,   // three-term Chebyshev: y (holding x_prev on entry, or prev holds it; not read
                    //   if cheb_c == 0) <- u + cheb_a * (dinv (b - A u) + cheb_c * (u - x_prev))
  MODE_CHEB3D = 6,  // CHEB3, then partial[t] = sum of b . y and partial[n_tiles + t] = sum of
                    //   b . (y - zold) over tile t's DOFs (the Polak-Ribiere dots of the last step)
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
  double* partial;        // per-tile FP64 partial sums (DOT: u.Au; RESID: |y|^2), length >= n_tiles
  // load-time combination, the combined input written to u_out at the owned nodes:
  // LOAD_AXPY u <- u + s w (s = sdev[sidx], a device scalar; DOT and RESID passes of the Krylov
  // loops), LOAD_MUL u <- load_scale * u (.) w (CHEB3: the first Chebyshev step from zero)
  const void* w;
  const double* sdev;
  int sidx;
  T* u_out;
  float* y32;             // RESID: FP32 copy of the output (the V-cycle right-hand side)
  const T* prev;          // CHEB3 / CHEB3D: the previous iterate (default: y on entry)
  const T* zold;          // CHEB3D: the vector the new output is compared with
  int nx, ny, nz;
  int lx;                 // x chunk length (0: automatic)
  int tile;               // tile variant (MV_TILE_*)
  int cube;               // coef holds the 8 cube-pattern values instead of 72 block entries
  int flags;              // MV_PREV_ZERO: CHEB3's previous iterate is zero (y not read); MV_W_F32
  int fixed_zero;         // with a mask: the output is 0 at fixed DOFs (M K M u, (b - A u) M)
  int tiles_y, tiles_z, chunks_x, n_tiles;
  int tile_ty;            // owned node rows per y tile (at most the tile variant's TY)
  T omega, cheb_a, cheb_c, load_scale;
  T kdi[3];               // 1 / (diagonal of ke per component, the same at every corner of a brick):
                          // with dinv == nullptr, D^-1 = kdi[c] / (sum of the incident moduli)
  T coef[72];             // parity blocks C[g][i][j], or [p-q, q, a, b, c, d, e-f, f] (cube)
};

// Launch plan of the product for a grid: tile variant, x chunking and grid size, chosen from
// the CU count and the resident workgroups per CU of the kernel (hint values > 0 override).
struct MatvecPlan {
  int tile, lx, tiles_y, tiles_z, chunks_x, n_tiles, grid, tile_ty;
  double cost;
};
MatvecPlan matvec_plan_f64(bool cube, int nx, int ny, int nz, int lx_hint, int tile_hint);
MatvecPlan matvec_plan_f32(int storage, bool cube, int nx, int ny, int nz, int lx_hint, int tile_hint);
inline MatvecPlan matvec_plan(int dtype, int storage, bool cube, int nx, int ny, int nz, int lx_hint, int tile_hint) {
  return dtype == F64 ? matvec_plan_f64(cube, nx, ny, nz, lx_hint, tile_hint)
                      : matvec_plan_f32(storage, cube, nx, ny, nz, lx_hint, tile_hint);
}
bool cube_pattern(const double* blocks72, double* k8);

template <typename T>
int matvec_launch(MatvecParams<T> p, int storage, int mode, int mask_kind, hipStream_t st);

// A batch of products in one launch: descs (device) holds one MatvecParams<T> per problem with
// its plan filled in, offs (device, n + 1 entries) the prefix sums of the problems' tile counts.
// Every problem shares the mode, the mask kind, the tile variant, the load-time combination and
// the voxel-cube coefficient form.
struct MatvecBatch {
  int mode, mask_kind, tile, combine;   // combine: 0 none, 1 AXPY / MUL, 2 AXPY with an FP32 w (FP64)
  int n, total;
  const void* descs;
  const int* offs;
};
template <typename T> int matvec_batch_launch(const MatvecBatch& b, hipStream_t st);
template <typename T> MatvecPlan matvec_batch_plan(int nx, int ny, int nz, int lx_hint, int tile);
// x chunk lengths lx[i] of the n problems of one batched launch, chosen for the batch as a whole
template <typename T> void matvec_batch_chunks(int n, const int* nx, const int* ny, const int* nz, int tile, int* lx);

}  // namespace vh
