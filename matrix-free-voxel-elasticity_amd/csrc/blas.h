#pragma once
#include "common.h"

namespace vh {

constexpr int BLAS_BLOCK = 256;
constexpr int BLAS_MAX_GRID = 2048;     // 8 blocks per CU on 256 CUs; partial arrays are this long

enum RedKind : int { RED_DOT = 0, RED_DOT2 = 1, RED_DOTDIFF = 2, RED_NORM2 = 3 };
enum ScalarOp : int { OP_NONE = 0, OP_ALPHA = 1, OP_BETA_FR = 2, OP_BETA_PR = 3, OP_SET_RZ = 4 };
// Device scalar slots (FP64).  Slots >= 8 are free for callers (norms, dots).
enum Slot : int { S_RZ = 0, S_PQ = 1, S_ALPHA = 2, S_BETA = 3, S_FLAG = 4, S_RZNEW = 5, S_NUM = 6, S_NORM2 = 7 };

int blas_grid(size_t n);
template <typename T> int reduce_launch(int kind, size_t n, const T* x, const T* y, const T* w, double* partial, hipStream_t st);
int finish_launch(const double* partial, int nparts, int count, double* s, int dst, int op, hipStream_t st);
template <typename T> int cg_update_launch(size_t n, T* x, const T* p, T* r, const T* q, const T* dinv, T* z,
                                           const double* s, int update_r, double* partial, hipStream_t st);
template <typename T> int jacobi_dot_launch(size_t n, const T* r, const T* dinv, T* z, double* partial, hipStream_t st);
template <typename T> int xpby_launch(size_t n, const T* z, T* p, const double* s, T* zcopy, hipStream_t st);
template <typename T> int axpby_launch(size_t n, T a, const T* x, const T* w, T b, T* y, hipStream_t st);
template <typename Ti, typename To> int convert_launch(size_t n, const Ti* x, To* y, double scale, const double* dscale,
                                                       int inv_scale, int accumulate, hipStream_t st);
template <typename Ti, typename Tv>
int cvt_dotdiff_launch(size_t n, const Ti* zin, const Tv* r, const Tv* zold, Tv* znew, double* partial, hipStream_t st);
template <typename TA, typename TV>
int dense_mv_launch(int n, int ld, const TA* A, const TV* r, TV* y, const uint8_t* mask, hipStream_t st);

}  // namespace vh
