#pragma once
#include "common.h"

namespace vh {

constexpr int CONE_MAX_OFF = 125;    // supports radii < 3 (support (2R+1)^3, R = ceil(r) - 1 <= 2)

template <typename T>
struct ConeArgs {
  int nx, ny, nz, n_off;
  int dx[CONE_MAX_OFF], dy[CONE_MAX_OFF], dz[CONE_MAX_OFF];
  T w[CONE_MAX_OFF];
};

template <typename T> int cone_launch(const T* x, const T* pre_div, const T* post_div, T* y, const ConeArgs<T>& c,
                                      hipStream_t st);
template <typename T> int oc_probe_launch(size_t n, const T* x, const T* dc, const T* dv, const T* w, T lmid, T move,
                                          T* xnew, double* partial, int nparts, hipStream_t st);
// up to OC_MAX_PROBES bisection multipliers evaluated in one pass (oc_probe_multi_kernel)
constexpr int OC_MAX_PROBES = 16;
struct OcLambdas {
  int count;
  double v[OC_MAX_PROBES];
};
template <typename T> int oc_probe_multi_launch(size_t n, const T* x, const T* dc, const T* dv, const T* w,
                                                const OcLambdas& lm, T move, double* partial, int nparts,
                                                hipStream_t st);

int heaviside_launch(size_t n, const double* rb, double beta, double eta, double tb, double den, double penal,
                     double E0, double Emin, double* rho, double* drho, double* E, double* partial, int nparts,
                     hipStream_t st);
int oc_design_launch(size_t n, const double* rho, const double* dc, const double* dv, double lmid, double move,
                     double lb, double* rho_new, hipStream_t st);

}  // namespace vh
