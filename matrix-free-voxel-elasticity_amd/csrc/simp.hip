// SIMP kernels around the solve: the cone density filter and the OC bisection probe.
//
// Filter (experiments/session_a/donor/src/gpu_fem/filter_r2.py:40-183): y_e = sum over the
// in-grid neighbours e+d with |d| < r of (r - |d|) x_{e+d}; forward = H x / s, transpose =
// H (g / s) (H is symmetric), s = H 1.  The offset list and weights are passed from the host.
//
// OC probe (experiments/session_h/run_w1b.py:46-62, 82-94): for a trial multiplier l,
// x_new = clip(x sqrt(max(-dc/dv/l, 0)), max(x - move, 0), min(x + move, 1)) and the filtered
// volume mean(H x_new / s).  The square root is split as (x sqrt(max(-dc/dv, 0))) l^(-1/2), the
// first factor formed once per element and l^(-1/2) once per probe: a probe then costs one
// multiply and the clamps instead of an IEEE divide and square root (values within 2 ulp of the
// unsplit form; the single- and multi-probe kernels evaluate the same expression).  Because H is symmetric, mean(H x_new / s) = sum_j x_new_j w_j / n
// with w = H (1/s), so one probe is a single fused elementwise pass plus a deterministic
// reduction instead of a filter application.
#include <algorithm>

#include "common.h"
#include "simp.h"

namespace vh {

template <typename T>
__global__ void __launch_bounds__(256) cone_kernel(const T* __restrict__ x, const T* __restrict__ pre_div,
                                                  const T* __restrict__ post_div, T* __restrict__ y,
                                                  const ConeArgs<T> c) {
  const size_t e = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  const int nx = c.nx, ny = c.ny, nz = c.nz;
  if (e >= size_t(nx) * ny * nz) return;
  const int ez = int(e % nz), ey = int((e / nz) % ny), ex = int(e / (size_t(nz) * ny));
  T acc = T(0);
  for (int k = 0; k < c.n_off; ++k) {
    const int jx = ex + c.dx[k], jy = ey + c.dy[k], jz = ez + c.dz[k];
    if (jx < 0 || jx >= nx || jy < 0 || jy >= ny || jz < 0 || jz >= nz) continue;
    const size_t j = (size_t(jx) * ny + jy) * nz + jz;
    T v = x[j];
    if (pre_div) v = v / pre_div[j];
    acc += c.w[k] * v;
  }
  if (post_div) acc = acc / post_div[e];
  y[e] = acc;
}

template <typename T>
__global__ void __launch_bounds__(256) oc_probe_kernel(size_t n, const T* __restrict__ x, const T* __restrict__ dc,
                                                      const T* __restrict__ dv, const T* __restrict__ w, T lmid,
                                                      T move, T* __restrict__ xnew, double* __restrict__ partial) {
  __shared__ double red[4];
  double acc = 0.0;
  const T rl = T(1.0 / sqrt(double(lmid)));
  for (size_t i = blockIdx.x * size_t(256) + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
    const T xi = x[i];
    const T q = -dc[i] / (dv ? dv[i] : T(1));
    T v = (xi * sqrt(q > T(0) ? q : T(0))) * rl;
    const T lo = xi - move > T(0) ? xi - move : T(0);
    const T hi = xi + move < T(1) ? xi + move : T(1);
    v = v < lo ? lo : (v > hi ? hi : v);
    if (xnew) xnew[i] = v;
    acc += double(v) * double(w ? w[i] : T(1));
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Several bisection probes in one pass over the design: partial[k * nparts + block] is the block's
// sum of xnew(lmid_k) * w, the per-element value and the in-block sum order exactly those of
// oc_probe_kernel (a probe's partials are bitwise the single probe's); xnew is not written.
template <typename T>
__global__ void __launch_bounds__(256) oc_probe_multi_kernel(size_t n, const T* __restrict__ x,
                                                            const T* __restrict__ dc, const T* __restrict__ dv,
                                                            const T* __restrict__ w, OcLambdas lm, T move,
                                                            double* __restrict__ partial) {
  __shared__ double red[OC_MAX_PROBES][4];
  double acc[OC_MAX_PROBES];
  T rl[OC_MAX_PROBES];
#pragma unroll
  for (int k = 0; k < OC_MAX_PROBES; ++k) {
    acc[k] = 0.0;
    rl[k] = k < lm.count ? T(1.0 / sqrt(double(T(lm.v[k])))) : T(0);   // (l rounded as the single probe's)
  }
  for (size_t i = blockIdx.x * size_t(256) + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
    const T xi = x[i];
    const T q = -dc[i] / (dv ? dv[i] : T(1));
    const T xs = xi * sqrt(q > T(0) ? q : T(0));
    const T lo = xi - move > T(0) ? xi - move : T(0);
    const T hi = xi + move < T(1) ? xi + move : T(1);
    const double wi = double(w ? w[i] : T(1));
#pragma unroll
    for (int k = 0; k < OC_MAX_PROBES; ++k) {
      if (k < lm.count) {
        T v = xs * rl[k];
        v = v < lo ? lo : (v > hi ? hi : v);
        acc[k] += double(v) * wi;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < OC_MAX_PROBES; ++k) {
    if (k < lm.count) {
      const double a = wave_sum(acc[k]);
      if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = a;
    }
  }
  __syncthreads();
  if (threadIdx.x < lm.count) {
    const int k = threadIdx.x;
    partial[size_t(k) * gridDim.x + blockIdx.x] = red[k][0] + red[k][1] + red[k][2] + red[k][3];
  }
}

// Heaviside projection of the filtered density (simp_r2.py:388-404) with its derivative, the
// SIMP modulus of the projected field and per-block sums of the projected field (its volume):
//   rho = (tb + tanh(beta (rb - eta))) / den,  drho/drb = beta / (den cosh^2(beta (rb - eta))),
//   tb = tanh(beta eta), den = tb + tanh(beta (1 - eta)) (host), identity for beta <= 1e-12;
//   E = Emin + (E0 - Emin) rho^penal.  Every output is optional.
__global__ void __launch_bounds__(256) heaviside_kernel(size_t n, const double* __restrict__ rb, double beta,
                                                       double eta, double tb, double den, double penal, double E0,
                                                       double Emin, double* __restrict__ rho, double* __restrict__ drho,
                                                       double* __restrict__ E, double* __restrict__ partial) {
  __shared__ double red[4];
  double acc = 0.0;
  const bool identity = beta <= 1e-12;
  for (size_t i = blockIdx.x * size_t(256) + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
    const double x = rb[i];
    double r = x, d = 1.0;
    if (!identity) {
      const double t = beta * (x - eta);
      r = (tb + tanh(t)) / den;
      const double ch = cosh(t);
      d = beta / (den * ch * ch);
    }
    if (rho) rho[i] = r;
    if (drho) drho[i] = d;
    if (E) E[i] = Emin + (E0 - Emin) * pow(r, penal);
    acc += r;
  }
  if (partial) {
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
  }
}

// Candidate design of one OC bisection probe on the projected formulation (simp_r2.py:497-507):
// rho_new = clip(clip(rho sqrt(-min(dc, -1e-30) / (lmid max(dv, 1e-30))), rho - move, rho + move), lb, 1).
__global__ void __launch_bounds__(256) oc_design_kernel(size_t n, const double* __restrict__ rho,
                                                       const double* __restrict__ dc, const double* __restrict__ dv,
                                                       double lmid, double move, double lb,
                                                       double* __restrict__ rho_new) {
  for (size_t i = blockIdx.x * size_t(256) + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
    const double x = rho[i];
    const double dcn = dc[i] < -1e-30 ? dc[i] : -1e-30;
    const double dvp = dv[i] > 1e-30 ? dv[i] : 1e-30;
    double v = x * sqrt(-dcn / (lmid * dvp));
    v = v < x - move ? x - move : (v > x + move ? x + move : v);
    v = v < lb ? lb : (v > 1.0 ? 1.0 : v);
    rho_new[i] = v;
  }
}

int heaviside_launch(size_t n, const double* rb, double beta, double eta, double tb, double den, double penal,
                     double E0, double Emin, double* rho, double* drho, double* E, double* partial, int nparts,
                     hipStream_t st) {
  hipLaunchKernelGGL(heaviside_kernel, dim3(nparts), dim3(256), 0, st, n, rb, beta, eta, tb, den, penal, E0, Emin, rho,
                     drho, E, partial);
  return hip_status(hipGetLastError(), "heaviside_kernel");
}

int oc_design_launch(size_t n, const double* rho, const double* dc, const double* dv, double lmid, double move,
                     double lb, double* rho_new, hipStream_t st) {
  const int grid = int(std::min<size_t>((n + 255) / 256, 4096));
  hipLaunchKernelGGL(oc_design_kernel, dim3(grid), dim3(256), 0, st, n, rho, dc, dv, lmid, move, lb, rho_new);
  return hip_status(hipGetLastError(), "oc_design_kernel");
}

template <typename T>
int cone_launch(const T* x, const T* pre_div, const T* post_div, T* y, const ConeArgs<T>& c, hipStream_t st) {
  const size_t ne = size_t(c.nx) * c.ny * c.nz;
  hipLaunchKernelGGL(cone_kernel<T>, dim3(int((ne + 255) / 256)), dim3(256), 0, st, x, pre_div, post_div, y, c);
  return hip_status(hipGetLastError(), "cone_kernel");
}

template <typename T>
int oc_probe_launch(size_t n, const T* x, const T* dc, const T* dv, const T* w, T lmid, T move, T* xnew,
                    double* partial, int nparts, hipStream_t st) {
  hipLaunchKernelGGL(oc_probe_kernel<T>, dim3(nparts), dim3(256), 0, st, n, x, dc, dv, w, lmid, move, xnew, partial);
  return hip_status(hipGetLastError(), "oc_probe_kernel");
}

template <typename T>
int oc_probe_multi_launch(size_t n, const T* x, const T* dc, const T* dv, const T* w, const OcLambdas& lm, T move,
                          double* partial, int nparts, hipStream_t st) {
  hipLaunchKernelGGL(oc_probe_multi_kernel<T>, dim3(nparts), dim3(256), 0, st, n, x, dc, dv, w, lm, move, partial);
  return hip_status(hipGetLastError(), "oc_probe_multi_kernel");
}

template int oc_probe_multi_launch<double>(size_t, const double*, const double*, const double*, const double*,
                                           const OcLambdas&, double, double*, int, hipStream_t);
template int oc_probe_multi_launch<float>(size_t, const float*, const float*, const float*, const float*,
                                          const OcLambdas&, float, double*, int, hipStream_t);
template int cone_launch<double>(const double*, const double*, const double*, double*, const ConeArgs<double>&, hipStream_t);
template int cone_launch<float>(const float*, const float*, const float*, float*, const ConeArgs<float>&, hipStream_t);
template int oc_probe_launch<double>(size_t, const double*, const double*, const double*, const double*, double, double,
                                     double*, double*, int, hipStream_t);
template int oc_probe_launch<float>(size_t, const float*, const float*, const float*, const float*, float, float, float*,
                                    double*, int, hipStream_t);

}  // namespace vh
