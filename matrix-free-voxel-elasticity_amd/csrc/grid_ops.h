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

template <typename T> int diag_launch(const T* E, const GridArgs<T>& g, T* out, int invert, hipStream_t st);
template <typename T> int energy_launch(const T* u, const GridArgs<T>& g, T* ce, const T* rho, T* dc, T penal, T dE, hipStream_t st);
template <typename T> int dense_launch(const T* u, const T* E, const T* ke, T* y, int nx, int ny, int nz, hipStream_t st);
template <typename T> int restrict_launch(const T* rf, T* rc, const GridArgs<T>& g, const uint8_t* cmask, int cmask_kind, hipStream_t st);
template <typename T> int prolong_launch(const T* ec, T* xf, const GridArgs<T>& g, int assign, hipStream_t st);
template <typename T> int coarsen_modulus_launch(const T* Ef, T* Ec, int nx, int ny, int nz, hipStream_t st);

}  // namespace vh
