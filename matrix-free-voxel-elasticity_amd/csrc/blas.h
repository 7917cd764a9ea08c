#pragma once
#include "common.h"

namespace vh {

constexpr int BLAS_BLOCK = 256;
constexpr int BLAS_MAX_GRID = 2048;     // 8 blocks per CU on 256 CUs; partial arrays are this long

enum RedKind : int { RED_DOT = 0, RED_DOT2 = 1, RED_DOTDIFF = 2, RED_NORM2 = 3 };
enum ScalarOp : int { OP_NONE = 0, OP_ALPHA = 1, OP_BETA_FR = 2, OP_BETA_PR = 3, OP_SET_RZ = 4, OP_INVSQRT = 5,
                      OP_RESNORM = 6 };
// Device scalar slots (FP64).  Slots 8, 9, 13.. are free for callers (norms, dots); S_DONE != 0
// freezes a Krylov recurrence (alpha = beta = 0: the iterate stops moving), which the batched
// solvers set for problems that have converged while others iterate on, and OP_RESNORM sets
// once the squared residual norm reaches S_TOL2 (S_ITS counts the iterations up to then).
enum Slot : int { S_RZ = 0, S_PQ = 1, S_ALPHA = 2, S_BETA = 3, S_FLAG = 4, S_RZNEW = 5, S_NUM = 6, S_NORM2 = 7,
                  S_DONE = 10, S_TOL2 = 11, S_ITS = 12 };

// Stores the `count` reduced values v into s[dst..] and evaluates the scalar recurrence `op`.
// With OP_NONE only s[dst .. dst + count) is touched (s may then be a single output value);
// the recurrences read and write the slots of a 16-value scalar block.
__device__ inline void apply_scalar_op(double* s, const double* v, int count, int dst, int op) {
  for (int j = 0; j < count; ++j) s[dst + j] = v[j];
  if (op == OP_NONE) return;
  const bool done = s[S_DONE] != 0.0;
  switch (op) {
    case OP_ALPHA: {         // alpha = rz / pq  (pq just written to s[dst])
      const double pq = v[0];
      if (done) s[S_ALPHA] = 0.0;
      else if (pq == 0.0 || !isfinite(pq)) { s[S_ALPHA] = 0.0; s[S_FLAG] = 1.0; }
      else s[S_ALPHA] = s[S_RZ] / pq;
      break;
    }
    case OP_BETA_FR: {       // Fletcher-Reeves: beta = rz_new / rz ; rz = rz_new
      const double rzn = v[0];
      const double rz = s[S_RZ];
      s[S_BETA] = (rz == 0.0 || done) ? 0.0 : rzn / rz;
      if (rz == 0.0) s[S_FLAG] = 1.0;
      s[S_RZ] = rzn;
      break;
    }
    case OP_BETA_PR: {       // Polak-Ribiere: beta = r_new.(z_new - z) / rz ; rz = r_new.z_new
      const double rzn = v[0], num = v[1];
      const double rz = s[S_RZ];
      s[S_BETA] = (rz == 0.0 || done) ? 0.0 : num / rz;
      if (rz == 0.0) s[S_FLAG] = 1.0;
      s[S_RZ] = rzn;
      break;
    }
    case OP_SET_RZ: s[S_RZ] = v[0]; break;
    case OP_INVSQRT: s[dst] = v[0] > 0.0 ? 1.0 / sqrt(v[0]) : 0.0; break;   // normalisation factor
    case OP_RESNORM:         // an iteration's squared residual norm (in s[dst]): count, test, freeze
      if (!done) {
        s[S_ITS] += 1.0;
        if (v[0] <= s[S_TOL2]) s[S_DONE] = 1.0;
      }
      break;
    default: break;
  }
}

int blas_grid(size_t n);
template <typename T> int reduce_launch(int kind, size_t n, const T* x, const T* y, const T* w, double* partial, hipStream_t st);
int finish_launch(const double* partial, int nparts, int count, double* s, int dst, int op, hipStream_t st);
template <typename T> int cg_update_launch(size_t n, T* x, const T* p, T* r, const T* q, const T* dinv, T* z,
                                           const double* s, int update_r, double* partial, hipStream_t st);
template <typename T> int jacobi_dot_launch(size_t n, const T* r, const T* dinv, T* z, double* partial, hipStream_t st);
template <typename T> int xpby_launch(size_t n, const T* z, T* p, const double* s, T* zcopy, hipStream_t st);
template <typename T> int axpby_launch(size_t n, T a, const T* x, const T* w, T b, T* y, hipStream_t st);
template <typename T> int fill_splitmix_launch(size_t n, unsigned long long seed, T* y, hipStream_t st);
template <typename Ti, typename To> int convert_launch(size_t n, const Ti* x, To* y, double scale, const double* dscale,
                                                       int inv_scale, int accumulate, hipStream_t st);
int convert_bf16_launch(size_t n, const float* x, uint16_t* y, double scale, const double* dscale, int inv_scale,
                        hipStream_t st);
template <typename Ti, typename Tv>
int cvt_dotdiff_launch(size_t n, const Ti* zin, const Tv* r, const Tv* zold, Tv* znew, double* partial, hipStream_t st);
template <typename TA, typename TV>
int dense_mv_launch(int n, int ld, const TA* A, const TV* r, TV* y, const uint8_t* mask, hipStream_t st);

}  // namespace vh
