// C-ABI entry points of libvoxelhip.so: argument checks, dtype dispatch, host-side setup.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "voxelhip.h"
#include "common.h"
#include "matvec.h"
#include "stencil.h"
#include "lanes.h"
#include "grid_ops.h"
#include "blas.h"
#include "simp.h"

namespace vh {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int hip_status(hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  return fail(-1, "%s: %s", what, hipGetErrorString(e));
}

// Modal algebra of the 8-node brick (binary lane l = x + 2y + 4z; native order of ke).
static const int kNativeLane[8] = {0, 1, 3, 2, 4, 5, 7, 6};

static int modal_blocks(const double* ke, double tol, double* blocks, double* discarded) {
  // full = T ke T^T / 64 with T[3p+c][3n+c] = (-1)^popcount(p & lane(n))
  double full[24][24], tmp[24][24];
  auto h = [](int p, int l) { return (__builtin_popcount(p & l) & 1) ? -1.0 : 1.0; };
  for (int i = 0; i < 24; ++i)      // tmp = T ke
    for (int j = 0; j < 24; ++j) {
      const int p = i / 3, c = i % 3;
      double s = 0.0;
      for (int n = 0; n < 8; ++n) s += h(p, kNativeLane[n]) * ke[(3 * n + c) * 24 + j];
      tmp[i][j] = s;
    }
  for (int i = 0; i < 24; ++i)      // full = tmp T^T / 64
    for (int j = 0; j < 24; ++j) {
      const int p = j / 3, c = j % 3;
      double s = 0.0;
      for (int n = 0; n < 8; ++n) s += tmp[i][3 * n + c] * h(p, kNativeLane[n]);
      full[i][j] = s / 64.0;
    }
  bool in_block[24][24] = {};
  for (int g = 0; g < 8; ++g) {
    const int idx[3] = {3 * (g ^ 1), 3 * (g ^ 2) + 1, 3 * (g ^ 4) + 2};
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        blocks[9 * g + 3 * a + b] = full[idx[a]][idx[b]];
        in_block[idx[a]][idx[b]] = true;
      }
  }
  double out = 0.0, all = 0.0;
  for (int i = 0; i < 24; ++i)
    for (int j = 0; j < 24; ++j) {
      all += full[i][j] * full[i][j];
      if (!in_block[i][j]) out += full[i][j] * full[i][j];
    }
  const double rel = all > 0 ? std::sqrt(out / all) : NAN;
  if (discarded) *discarded = rel;
  if (!std::isfinite(rel) || rel > tol)
    return fail(-3, "Element is outside the declared parity-block model: %.3e", rel);
  return 0;
}

// The parity blocks of a voxel cube (any edge length, isotropic) follow one pattern with 8
// distinct values: B0 = (p-q) I + q 11^T;  B1 = [0 0 0; 0 a b; 0 b a], B2 = [a 0 b; 0 0 0;
// b 0 a], B4 = [a b 0; b a 0; 0 0 0];  B3 = [c c 0; c c 0; 0 0 d], B5 = [c 0 c; 0 d 0; c 0 c],
// B6 = [d 0 0; 0 c c; 0 c c];  B7 = (e-f) I + f 11^T.  Returns true and the values
// [p-q, q, a, b, c, d, e-f, f] when `blocks` matches it to 1e-12 of its largest entry.
bool cube_pattern(const double* B, double* k8) {
  double mx = 0.0;
  for (int i = 0; i < 72; ++i) mx = std::fmax(mx, std::fabs(B[i]));
  if (!(mx > 0.0)) return false;
  const double p = B[0], q = B[1], a = B[9 + 4], b = B[9 + 5], c = B[27 + 0], d = B[27 + 8];
  const double e = B[63 + 0], f = B[63 + 1];
  const double want[72] = {
      p, q, q, q, p, q, q, q, p,      // B0
      0, 0, 0, 0, a, b, 0, b, a,      // B1
      a, 0, b, 0, 0, 0, b, 0, a,      // B2
      c, c, 0, c, c, 0, 0, 0, d,      // B3
      a, b, 0, b, a, 0, 0, 0, 0,      // B4
      c, 0, c, 0, d, 0, c, 0, c,      // B5
      d, 0, 0, 0, c, c, 0, c, c,      // B6
      e, f, f, f, e, f, f, f, e};     // B7
  for (int i = 0; i < 72; ++i)
    if (std::fabs(B[i] - want[i]) > 1e-12 * mx) return false;
  k8[0] = p - q; k8[1] = q; k8[2] = a; k8[3] = b; k8[4] = c; k8[5] = d; k8[6] = e - f; k8[7] = f;
  return true;
}

// every kernel indexes DOFs with 32-bit counts below 2^31; the z extent is capped at 2048 elements
static bool dims_ok(int nx, int ny, int nz) {
  return nx >= 1 && ny >= 1 && nz >= 1 && nz <= 2048 &&
         3.0 * (nx + 1.0) * (ny + 1.0) * (nz + 1.0) < 2147483648.0;
}
static bool dtype_ok(int d) { return d == F64 || d == F32; }

}  // namespace vh

using namespace vh;

#define VH_REQUIRE(cond, ...) \
  do {                        \
    if (!(cond)) return fail(-2, __VA_ARGS__); \
  } while (0)

template <typename T>
static int matvec_t(int storage, const double* blocks72, const void* E, const void* u, void* y, int nx, int ny, int nz,
                    int mask_kind, const uint8_t* mask, const vh_epilogue* ep, hipStream_t st) {
  MatvecParams<T> p{};
  p.u = u;
  p.E = static_cast<const T*>(E);
  p.y = static_cast<T*>(y);
  p.mask = mask;
  p.nx = nx; p.ny = ny; p.nz = nz;
  double k8[8];
  if (ep && ep->w && !cube_pattern(blocks72, k8))
    return fail(-2, "vh_matvec: the load-time combination is implemented for voxel-cube blocks");
  if (cube_pattern(blocks72, k8)) {
    p.cube = 1;
    for (int i = 0; i < 8; ++i) p.coef[i] = T(k8[i]);
  } else {
    for (int i = 0; i < 72; ++i) p.coef[i] = T(blocks72[i]);
  }
  int mode = MODE_RAW;
  if (ep) {
    mode = ep->mode;
    p.b = static_cast<const T*>(ep->b);
    p.dinv = static_cast<const T*>(ep->dinv);
    p.d = static_cast<T*>(ep->d);
    p.partial = ep->partial;
    p.omega = T(ep->omega);
    p.cheb_a = T(ep->cheb_a);
    p.cheb_c = T(ep->cheb_c);
    p.lx = ep->lx;
    p.tile = ep->tile;
    for (int c = 0; c < 3; ++c) p.kdi[c] = ep->kdiag[c] > 0 ? T(1.0 / ep->kdiag[c]) : T(0);
    p.flags = ep->flags;
    p.w = ep->w;
    p.sdev = ep->sdev;
    p.sidx = ep->sidx;
    p.u_out = static_cast<T*>(ep->u_out);
    p.y32 = static_cast<float*>(ep->y32);
    p.fixed_zero = ep->fixed_zero;
    p.load_scale = T(ep->load_scale);
    p.prev = static_cast<const T*>(ep->prev);
    p.zold = static_cast<const T*>(ep->zold);
  }
  return matvec_launch<T>(p, storage, mode, mask_kind, st);
}

template <typename T>
static GridArgs<T> grid_args(int nx, int ny, int nz, int mask_kind, const uint8_t* mask, const double* ke576,
                             const double* blocks72) {
  GridArgs<T> g{};
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.mask_kind = mask_kind;
  g.mask = mask;
  if (ke576)
    for (int i = 0; i < 24; ++i) g.kediag[i] = T(ke576[i * 24 + i]);
  if (blocks72)
    for (int i = 0; i < 72; ++i) g.coef[i] = T(blocks72[i]);
  return g;
}

template <typename T>
static int stencil_t(const void* U, const void* u, void* y, int nx, int ny, int nz, int mask_kind, const uint8_t* mask,
                     const vh_epilogue* ep, hipStream_t st) {
  StencilParams<T> p{};
  p.U = static_cast<const T*>(U);
  p.u = static_cast<const T*>(u);
  p.y = static_cast<T*>(y);
  p.mask = mask;
  p.nx = nx; p.ny = ny; p.nz = nz;
  int mode = MODE_RAW;
  if (ep) {
    mode = ep->mode;
    p.b = static_cast<const T*>(ep->b);
    p.dinv = static_cast<const T*>(ep->dinv);
    p.prev = static_cast<const T*>(ep->prev);
    p.y32 = ep->y32;
    p.fixed_zero = ep->fixed_zero;
    p.flags = ep->flags;
    p.omega = T(ep->omega);
    p.cheb_a = T(ep->cheb_a);
    p.cheb_c = T(ep->cheb_c);
  }
  return stencil_apply_launch<T>(p, mode, mask_kind, st);
}

// Device copies of 24x24 element matrices for the dense node kernel, one per (device, dtype,
// contents), kept for the life of the process: a call with a matrix seen before launches without
// a copy (and can be captured in a HIP graph); a new matrix is copied on the call's stream.
struct KeEntry {
  int device, dtype;
  double ke[576];
  void* ptr;
};
static std::vector<KeEntry>& ke_cache() {
  static std::vector<KeEntry> c;
  return c;
}
static std::mutex& ke_mutex() {
  static std::mutex m;
  return m;
}

static int device_ke(int dtype, const double* ke576, hipStream_t st, const void** out) {
  int dev = 0;
  if (int r = hip_status(hipGetDevice(&dev), "hipGetDevice")) return r;
  std::lock_guard<std::mutex> lock(ke_mutex());
  for (const KeEntry& e : ke_cache())
    if (e.device == dev && e.dtype == dtype && std::memcmp(e.ke, ke576, sizeof(e.ke)) == 0) {
      *out = e.ptr;
      return 0;
    }
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (int r = hip_status(hipStreamIsCapturing(st, &cap), "hipStreamIsCapturing")) return r;
  if (cap != hipStreamCaptureStatusNone)
    return fail(-2, "vh_dense_product: a new ke inside a graph capture (call once before capturing)");
  KeEntry e{};
  e.device = dev;
  e.dtype = dtype;
  std::memcpy(e.ke, ke576, sizeof(e.ke));
  const size_t bytes = 576 * (dtype == F64 ? sizeof(double) : sizeof(float));
  if (hipMalloc(&e.ptr, bytes) != hipSuccess) return fail(-1, "vh_dense_product: hipMalloc ke");
  float k32[576];
  for (int i = 0; i < 576; ++i) k32[i] = float(ke576[i]);
  int r = hip_status(hipMemcpyAsync(e.ptr, dtype == F64 ? static_cast<const void*>(ke576) : k32, bytes,
                                    hipMemcpyHostToDevice, st), "ke copy");
  if (!r) r = hip_status(hipStreamSynchronize(st), "ke copy");       // the host arrays may go away
  if (r) {
    (void)hipFree(e.ptr);
    return r;
  }
  ke_cache().push_back(e);
  *out = e.ptr;
  return 0;
}

// ---- allocation-free DOT pass: the per-tile partial sums live in a library-owned device array.
// Each GPU holds its own copy; host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device (the array's name is a host-side
// stand-in), looked up once per device on the first call there and kept in g_scratch_addr.
constexpr int kScratchTiles = 65536;     // 512 KiB of FP64 partial sums per device
__device__ __align__(256) double scratch[kScratchTiles];
constexpr int kMaxDevices = 64;
static double* g_scratch_addr[kMaxDevices] = {};
static std::mutex g_scratch_mutex;

static int scratch_address(double** out) {
  int dev = 0;
  if (int r = hip_status(hipGetDevice(&dev), "hipGetDevice")) return r;
  if (dev < 0 || dev >= kMaxDevices) return fail(-2, "vh_matvec_dot_noalloc: device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_addr[dev] == nullptr) {
    void* ptr = nullptr;
    if (int r = hip_status(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return r;
    if (ptr == nullptr) return fail(-1, "hipGetSymbolAddress(scratch): null address");
    g_scratch_addr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_addr[dev];
  return 0;
}

extern "C" {

int vh_abi_version(void) { return VH_ABI_VERSION; }
const char* vh_last_error(void) { return g_err; }

int vh_modal_blocks(const double* ke576, double tolerance, double* blocks72, double* discarded) {
  VH_REQUIRE(ke576 && blocks72, "vh_modal_blocks: null pointer");
  return modal_blocks(ke576, tolerance, blocks72, discarded);
}

int vh_blocks_are_cube(const double* blocks72) {
  double k8[8];
  return blocks72 != nullptr && cube_pattern(blocks72, k8) ? 1 : 0;
}

int vh_matvec_tiles(int dtype, int storage, const double* blocks72, int nx, int ny, int nz, int lx, int tile,
                    int* n_tiles, int* n_workgroups) {
  VH_REQUIRE(dims_ok(nx, ny, nz), "vh_matvec_tiles: bad grid %d x %d x %d", nx, ny, nz);
  VH_REQUIRE(dtype_ok(dtype), "vh_matvec_tiles: dtype %d", dtype);
  double k8[8];
  const bool cube = blocks72 == nullptr || cube_pattern(blocks72, k8);
  const MatvecPlan pl = matvec_plan(dtype, storage, cube, nx, ny, nz, lx, tile);
  if (n_tiles) *n_tiles = pl.n_tiles;
  if (n_workgroups) *n_workgroups = pl.grid;
  return 0;
}


int vh_matvec_fused(int dtype, int storage, const double* blocks72, const void* E, const void* u, void* y, int nx,
                    int ny, int nz, int mask_kind, const uint8_t* mask, const vh_epilogue* ep, void* stream) {
  VH_REQUIRE(dtype_ok(dtype), "vh_matvec: dtype %d not supported", dtype);
  VH_REQUIRE(storage == dtype || (storage == BF16 && dtype == F32) ||
                 (storage == F32 && dtype == F64 && ep && ep->mode == MODE_DOT && ep->w),
             "vh_matvec: storage %d with dtype %d", storage, dtype);
  VH_REQUIRE(blocks72 && E && u && (y || (ep && (ep->mode == MODE_DOT || (ep->mode == MODE_RESID && ep->y32)))),
             "vh_matvec: null pointer");
  VH_REQUIRE(u != y, "vh_matvec: y must not alias u");
  VH_REQUIRE(dims_ok(nx, ny, nz), "vh_matvec: bad grid %d x %d x %d", nx, ny, nz);
  VH_REQUIRE(mask_kind >= 0 && mask_kind <= 2, "vh_matvec: bad mask kind %d", mask_kind);
  VH_REQUIRE(mask_kind != MASK_ARRAY || mask, "vh_matvec: mask array missing");
  if (ep) {
    VH_REQUIRE(ep->mode >= 0 && ep->mode <= MODE_CHEB3D, "vh_matvec: bad mode %d", ep->mode);
    const int m = ep->mode;
    const bool smoother = m == MODE_JACOBI || m == MODE_CHEB || m == MODE_CHEB3 || m == MODE_CHEB3D;
    VH_REQUIRE(m != MODE_CHEB3D || (ep->partial && ep->zold && ep->zold != y), "vh_matvec: CHEB3D needs partial, zold");
    VH_REQUIRE(!ep->prev || ((m == MODE_CHEB3 || m == MODE_CHEB3D) && ep->prev != y),
               "vh_matvec: prev is the previous iterate of a CHEB3 pass (not y)");
    VH_REQUIRE(!(m == MODE_RESID || smoother) || ep->b, "vh_matvec: b missing");
    VH_REQUIRE(m != MODE_CHEB || ep->d, "vh_matvec: d missing");
    VH_REQUIRE(m != MODE_DOT || ep->partial, "vh_matvec: partial missing");
    VH_REQUIRE(!ep->w || (((m == MODE_DOT || m == MODE_RESID) && ep->sdev) || (m == MODE_CHEB3 && !ep->prev)),
               "vh_matvec: the load-time combination needs a DOT/RESID pass with sdev, or a CHEB3 pass");
    VH_REQUIRE(!ep->w || (ep->u_out && ep->u_out != u && ep->u_out != y && ep->w != y && ep->u_out != ep->w &&
                          ep->u_out != ep->b),
               "vh_matvec: the load-time combination needs a distinct u_out");
    VH_REQUIRE(!ep->w || m != MODE_CHEB3 || (ep->flags & MV_PREV_ZERO) || ep->cheb_c == 0.0,
               "vh_matvec: a CHEB3 pass with the load-time product is the first step from a zero iterate "
               "(VH_PREV_ZERO or cheb_c = 0)");
    VH_REQUIRE(!ep->y32 || m == MODE_RESID, "vh_matvec: y32 is an output of the RESID pass");
    VH_REQUIRE((ep->flags & ~(MV_PREV_ZERO | MV_W_F32)) == 0, "vh_matvec: unknown flags %d", ep->flags);
    VH_REQUIRE(!(ep->flags & MV_W_F32) || (dtype == F64 && storage == F64 && m == MODE_RESID && ep->w),
               "vh_matvec: an FP32 w is read by the FP64 RESID pass with the load-time combination only");
    VH_REQUIRE(!y || (ep->b != y && ep->dinv != y), "vh_matvec: y must not alias an input");
    if (smoother && ep->dinv == nullptr)   // D^-1 from the moduli and ke's diagonal
      VH_REQUIRE(ep->kdiag[0] > 0 && ep->kdiag[1] > 0 && ep->kdiag[2] > 0, "vh_matvec: kdiag missing");
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == F64) return matvec_t<double>(storage, blocks72, E, u, y, nx, ny, nz, mask_kind, mask, ep, st);
  return matvec_t<float>(storage, blocks72, E, u, y, nx, ny, nz, mask_kind, mask, ep, st);
}

int vh_matvec(int dtype, int storage, const double* blocks72, const void* E, const void* u, void* y, int nx, int ny,
              int nz, int mask_kind, const uint8_t* mask, void* stream) {
  return vh_matvec_fused(dtype, storage, blocks72, E, u, y, nx, ny, nz, mask_kind, mask, nullptr, stream);
}

int vh_matvec_dot_noalloc(int dtype, const double* blocks72, const void* E, const void* u, void* y, int nx, int ny,
                          int nz, int mask_kind, const uint8_t* mask, double* pq, void* stream) {
  VH_REQUIRE(dtype_ok(dtype), "vh_matvec_dot_noalloc: dtype %d not supported", dtype);
  VH_REQUIRE(blocks72 && pq, "vh_matvec_dot_noalloc: null pointer");
  VH_REQUIRE(dims_ok(nx, ny, nz), "vh_matvec_dot_noalloc: bad grid %d x %d x %d", nx, ny, nz);
  double k8[8];
  const MatvecPlan pl = matvec_plan(dtype, dtype, cube_pattern(blocks72, k8), nx, ny, nz, 0, 0);
  VH_REQUIRE(pl.n_tiles >= 1 && pl.n_tiles <= kScratchTiles,
             "vh_matvec_dot_noalloc: %d tiles exceed the %d partial sums of the library's array", pl.n_tiles,
             kScratchTiles);
  double* partial = nullptr;
  if (int r = scratch_address(&partial)) return r;
  vh_epilogue ep{};
  ep.mode = MODE_DOT;
  ep.partial = partial;
  // every one of the plan's n_tiles tiles writes its partial sum, so the array needs no clearing
  if (int r = vh_matvec_fused(dtype, dtype, blocks72, E, u, y, nx, ny, nz, mask_kind, mask, &ep, stream)) return r;
  return finish_launch(partial, pl.n_tiles, 1, pq, 0, OP_NONE, static_cast<hipStream_t>(stream));
}

int vh_dense_product(int dtype, const double* ke576, const void* E, const void* u, void* y, int nx, int ny, int nz,
                     void* stream) {
  VH_REQUIRE(dtype_ok(dtype), "vh_dense_product: dtype %d", dtype);
  VH_REQUIRE(ke576 && E && u && y && u != y, "vh_dense_product: bad pointers");
  VH_REQUIRE(dims_ok(nx, ny, nz), "vh_dense_product: bad grid");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const void* dk = nullptr;
  if (int r = device_ke(dtype, ke576, st, &dk)) return r;
  if (dtype == F64)
    return dense_launch<double>(static_cast<const double*>(u), static_cast<const double*>(E), static_cast<const double*>(dk),
                                static_cast<double*>(y), nx, ny, nz, st);
  return dense_launch<float>(static_cast<const float*>(u), static_cast<const float*>(E), static_cast<const float*>(dk),
                             static_cast<float*>(y), nx, ny, nz, st);
}


int vh_lanes24_product(int dtype, const double* blocks72, const void* E, const void* u, void* y, int nx, int ny,
                       int nz, void* stream) {
  VH_REQUIRE(dtype_ok(dtype), "vh_lanes24_product: dtype %d", dtype);
  VH_REQUIRE(blocks72 && E && u && y && u != y, "vh_lanes24_product: bad pointers");
  VH_REQUIRE(dims_ok(nx, ny, nz) && 32.0 * nx * ny * nz < 2147483648.0, "vh_lanes24_product: bad grid");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == F64)
    return lanes24_launch<double>(blocks72, static_cast<const double*>(E), static_cast<const double*>(u),
                                  static_cast<double*>(y), nx, ny, nz, st);
  return lanes24_launch<float>(blocks72, static_cast<const float*>(E), static_cast<const float*>(u),
                               static_cast<float*>(y), nx, ny, nz, st);
}

int vh_diagonal(int dtype, const double* ke576, const void* E, void* out, int nx, int ny, int nz, int mask_kind,
                const uint8_t* mask, int invert, void* stream) {
  VH_REQUIRE(dtype_ok(dtype) && ke576 && E && out && dims_ok(nx, ny, nz), "vh_diagonal: bad arguments");
  VH_REQUIRE(mask_kind >= 0 && mask_kind <= 2 && (mask_kind != MASK_ARRAY || mask), "vh_diagonal: bad mask");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == F64)
    return diag_launch<double>(static_cast<const double*>(E), grid_args<double>(nx, ny, nz, mask_kind, mask, ke576, nullptr),
                               static_cast<double*>(out), invert, st);
  return diag_launch<float>(static_cast<const float*>(E), grid_args<float>(nx, ny, nz, mask_kind, mask, ke576, nullptr),
                            static_cast<float*>(out), invert, st);
}

int vh_element_energy(int dtype, const double* blocks72, const void* u, void* ce, int nx, int ny, int nz, int mask_kind,
                      const uint8_t* mask, const void* rho, void* dc, double penal, double dE, void* stream) {
  VH_REQUIRE(dtype_ok(dtype) && blocks72 && u && ce && dims_ok(nx, ny, nz), "vh_element_energy: bad arguments");
  VH_REQUIRE(!dc || rho, "vh_element_energy: rho missing");
  VH_REQUIRE(mask_kind >= 0 && mask_kind <= 2 && (mask_kind != MASK_ARRAY || mask), "vh_element_energy: bad mask");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == F64)
    return energy_launch<double>(static_cast<const double*>(u), grid_args<double>(nx, ny, nz, mask_kind, mask, nullptr, blocks72),
                                 static_cast<double*>(ce), static_cast<const double*>(rho), static_cast<double*>(dc), penal, dE, st);
  return energy_launch<float>(static_cast<const float*>(u), grid_args<float>(nx, ny, nz, mask_kind, mask, nullptr, blocks72),
                              static_cast<float*>(ce), static_cast<const float*>(rho), static_cast<float*>(dc), float(penal),
                              float(dE), st);
}

int vh_restrict(int dtype, const void* rf, void* rc, int nx, int ny, int nz, int fmask_kind, const uint8_t* fmask,
                int cmask_kind, const uint8_t* cmask, void* stream) {
  VH_REQUIRE(dtype_ok(dtype) && rf && rc && rf != rc && dims_ok(nx, ny, nz), "vh_restrict: bad arguments");
  VH_REQUIRE(nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0, "vh_restrict: fine grid %d x %d x %d not even", nx, ny, nz);
  VH_REQUIRE((fmask_kind != MASK_ARRAY || fmask) && (cmask_kind != MASK_ARRAY || cmask), "vh_restrict: mask missing");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == F64)
    return restrict_launch<double>(static_cast<const double*>(rf), static_cast<double*>(rc),
                                   grid_args<double>(nx, ny, nz, fmask_kind, fmask, nullptr, nullptr), cmask, cmask_kind, st);
  return restrict_launch<float>(static_cast<const float*>(rf), static_cast<float*>(rc),
                                grid_args<float>(nx, ny, nz, fmask_kind, fmask, nullptr, nullptr), cmask, cmask_kind, st);
}

int vh_prolong(int dtype, const void* ec, void* xf, int nx, int ny, int nz, int mask_kind, const uint8_t* mask,
               int assign, void* stream) {
  VH_REQUIRE(dtype_ok(dtype) && ec && xf && ec != xf && dims_ok(nx, ny, nz), "vh_prolong: bad arguments");
  VH_REQUIRE(nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0, "vh_prolong: fine grid %d x %d x %d not even", nx, ny, nz);
  VH_REQUIRE(mask_kind != MASK_ARRAY || mask, "vh_prolong: mask missing");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == F64)
    return prolong_launch<double>(static_cast<const double*>(ec), static_cast<double*>(xf),
                                  grid_args<double>(nx, ny, nz, mask_kind, mask, nullptr, nullptr), assign, st);
  return prolong_launch<float>(static_cast<const float*>(ec), static_cast<float*>(xf),
                               grid_args<float>(nx, ny, nz, mask_kind, mask, nullptr, nullptr), assign, st);
}

int vh_coarsen_modulus(int dtype, const void* Ef, void* Ec, int nx, int ny, int nz, void* stream) {
  VH_REQUIRE(dtype_ok(dtype) && Ef && Ec && Ef != Ec && dims_ok(nx, ny, nz), "vh_coarsen_modulus: bad arguments");
  VH_REQUIRE(nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0, "vh_coarsen_modulus: grid not even");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == F64)
    return coarsen_modulus_launch<double>(static_cast<const double*>(Ef), static_cast<double*>(Ec), nx, ny, nz, st);
  return coarsen_modulus_launch<float>(static_cast<const float*>(Ef), static_cast<float*>(Ec), nx, ny, nz, st);
}

int vh_assemble_dense(const double* ke, const double* E, int nx, int ny, int nz, int mask_kind, const uint8_t* mask,
                      double* A, int ld, void* stream) {
  VH_REQUIRE(ke && E && A && dims_ok(nx, ny, nz), "vh_assemble_dense: bad arguments");
  const double n = 3.0 * (nx + 1.0) * (ny + 1.0) * (nz + 1.0);
  VH_REQUIRE(n * n < 4.0e9 && ld >= n, "vh_assemble_dense: %.0f DOFs or ld %d out of range", n, ld);
  VH_REQUIRE(mask_kind != MASK_ARRAY || mask, "vh_assemble_dense: mask missing");
  return assemble_dense_launch(ke, E, nx, ny, nz, mask_kind, mask, A, ld, static_cast<hipStream_t>(stream));
}

int vh_galerkin_scatter(const double* G, int nx, int ny, int nz, int mask_kind, const uint8_t* mask, double* A, int ld,
                        void* stream) {
  VH_REQUIRE(G && A && G != A && dims_ok(nx, ny, nz), "vh_galerkin_scatter: bad arguments");
  const double n = 3.0 * (nx + 1.0) * (ny + 1.0) * (nz + 1.0);
  VH_REQUIRE(81.0 * n < 2.0e9 && ld >= n, "vh_galerkin_scatter: %.0f DOFs or ld %d out of range", n, ld);
  VH_REQUIRE(mask_kind != MASK_ARRAY || mask, "vh_galerkin_scatter: mask missing");
  return galerkin_scatter_launch(G, nx, ny, nz, mask_kind, mask, A, ld, static_cast<hipStream_t>(stream));
}

int vh_stencil_scatter(int dtype, const double* G, void* U, int nx, int ny, int nz, void* stream) {
  VH_REQUIRE(dtype_ok(dtype), "vh_stencil_scatter: dtype %d", dtype);
  VH_REQUIRE(G && U && static_cast<const void*>(G) != U && dims_ok(nx, ny, nz), "vh_stencil_scatter: bad arguments");
  const double nn = (nx + 1.0) * (ny + 1.0) * (nz + 1.0);
  VH_REQUIRE(126.0 * nn < 2.0e9, "vh_stencil_scatter: %.0f nodes out of range", nn);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == F64) return stencil_scatter_launch<double>(G, static_cast<double*>(U), nx, ny, nz, st);
  return stencil_scatter_launch<float>(G, static_cast<float*>(U), nx, ny, nz, st);
}

int vh_stencil_apply(int dtype, const void* U, const void* u, void* y, int nx, int ny, int nz, int mask_kind,
                     const uint8_t* mask, const vh_epilogue* ep, void* stream) {
  VH_REQUIRE(dtype_ok(dtype), "vh_stencil_apply: dtype %d", dtype);
  VH_REQUIRE(U && u && (y || (ep && ep->mode == MODE_RESID && ep->y32)), "vh_stencil_apply: null pointer");
  VH_REQUIRE(u != y && U != y, "vh_stencil_apply: y must not alias u or U");
  VH_REQUIRE(dims_ok(nx, ny, nz) && 126.0 * (nx + 1.0) * (ny + 1.0) * (nz + 1.0) < 2.0e9,
             "vh_stencil_apply: bad grid %d x %d x %d", nx, ny, nz);
  VH_REQUIRE(mask_kind >= 0 && mask_kind <= 2, "vh_stencil_apply: bad mask kind %d", mask_kind);
  VH_REQUIRE(mask_kind != MASK_ARRAY || mask, "vh_stencil_apply: mask array missing");
  if (ep) {
    const int m = ep->mode;
    VH_REQUIRE(m == MODE_RAW || m == MODE_RESID || m == MODE_JACOBI || m == MODE_CHEB3,
               "vh_stencil_apply: mode %d (RAW, RESID, JACOBI and CHEB3 are available)", m);
    VH_REQUIRE(!ep->d && !ep->partial && !ep->w && !ep->zold && !ep->u_out,
               "vh_stencil_apply: d, partial, w, zold and u_out are parity-product fields");
    VH_REQUIRE(m == MODE_RAW || ep->b, "vh_stencil_apply: b missing");
    VH_REQUIRE(!(m == MODE_JACOBI || m == MODE_CHEB3) || ep->dinv, "vh_stencil_apply: dinv missing");
    VH_REQUIRE(!ep->prev || (m == MODE_CHEB3 && ep->prev != y), "vh_stencil_apply: prev is a CHEB3 input");
    VH_REQUIRE(!ep->y32 || m == MODE_RESID, "vh_stencil_apply: y32 is an output of the RESID pass");
    VH_REQUIRE(!y || (ep->b != y && ep->dinv != y), "vh_stencil_apply: y must not alias an input");
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == F64) return stencil_t<double>(U, u, y, nx, ny, nz, mask_kind, mask, ep, st);
  return stencil_t<float>(U, u, y, nx, ny, nz, mask_kind, mask, ep, st);
}

int vh_dense_mv(int a_dtype, int v_dtype, int n, int ld, const void* A, const void* r, void* y, const uint8_t* mask,
                void* stream) {
  VH_REQUIRE(dtype_ok(a_dtype) && dtype_ok(v_dtype) && n >= 1 && ld >= n && ld % 4 == 0 && A && r && y && r != y,
             "vh_dense_mv: bad arguments");
  VH_REQUIRE((reinterpret_cast<uintptr_t>(A) & 15) == 0, "vh_dense_mv: A must be 16-byte aligned");
  VH_REQUIRE(!(a_dtype == F32 && v_dtype == F64), "vh_dense_mv: an FP32 matrix takes FP32 vectors");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (a_dtype == F64 && v_dtype == F64)
    return dense_mv_launch<double, double>(n, ld, static_cast<const double*>(A), static_cast<const double*>(r),
                                           static_cast<double*>(y), mask, st);
  if (a_dtype == F64)
    return dense_mv_launch<double, float>(n, ld, static_cast<const double*>(A), static_cast<const float*>(r),
                                          static_cast<float*>(y), mask, st);
  return dense_mv_launch<float, float>(n, ld, static_cast<const float*>(A), static_cast<const float*>(r),
                                       static_cast<float*>(y), mask, st);
}

int vh_reduce_parts(size_t n) { return blas_grid(n); }

int vh_reduce(int dtype, int kind, size_t n, const void* x, const void* y, const void* w, double* partial,
              void* stream) {
  VH_REQUIRE(dtype_ok(dtype) && x && partial && n >= 1, "vh_reduce: bad arguments");
  VH_REQUIRE(kind >= 0 && kind <= 3, "vh_reduce: bad kind");
  VH_REQUIRE(kind == RED_NORM2 || y, "vh_reduce: y missing");
  VH_REQUIRE(!(kind == RED_DOT2 || kind == RED_DOTDIFF) || w, "vh_reduce: w missing");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == F64)
    return reduce_launch<double>(kind, n, static_cast<const double*>(x), static_cast<const double*>(y),
                                 static_cast<const double*>(w), partial, st);
  return reduce_launch<float>(kind, n, static_cast<const float*>(x), static_cast<const float*>(y),
                              static_cast<const float*>(w), partial, st);
}

int vh_finish(const double* partial, int nparts, int count, double* scal, int dst, int op, void* stream) {
  VH_REQUIRE(partial && scal && nparts >= 1 && count >= 1 && count <= 2 && dst >= 0 && op >= OP_NONE &&
                 op <= OP_RESNORM, "vh_finish: bad arguments");
  return finish_launch(partial, nparts, count, scal, dst, op, static_cast<hipStream_t>(stream));
}

int vh_cg_update(int dtype, size_t n, void* x, const void* p, void* r, const void* q, const void* dinv, void* z,
                 const double* scal, int update_r, double* partial, void* stream) {
  VH_REQUIRE(dtype_ok(dtype) && x && p && scal && n >= 1, "vh_cg_update: bad arguments");
  VH_REQUIRE(!update_r || (r && q), "vh_cg_update: r/q missing");
  VH_REQUIRE(!(update_r && dinv) || (z && partial), "vh_cg_update: z/partial missing");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == F64)
    return cg_update_launch<double>(n, static_cast<double*>(x), static_cast<const double*>(p), static_cast<double*>(r),
                                    static_cast<const double*>(q), static_cast<const double*>(dinv), static_cast<double*>(z),
                                    scal, update_r, partial, st);
  return cg_update_launch<float>(n, static_cast<float*>(x), static_cast<const float*>(p), static_cast<float*>(r),
                                 static_cast<const float*>(q), static_cast<const float*>(dinv), static_cast<float*>(z), scal,
                                 update_r, partial, st);
}

int vh_cvt_dotdiff(int dtype_in, int dtype_v, size_t n, const void* zin, const void* r, const void* zold, void* znew,
                   double* partial, void* stream) {
  VH_REQUIRE(dtype_ok(dtype_in) && dtype_ok(dtype_v) && zin && r && zold && znew && partial && n >= 1,
             "vh_cvt_dotdiff: bad arguments");
  VH_REQUIRE(znew != zin && znew != zold && znew != r, "vh_cvt_dotdiff: znew aliases an input");
  VH_REQUIRE(!(dtype_in == F64 && dtype_v == F32), "vh_cvt_dotdiff: FP64 input into FP32 vectors");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype_v == F32)
    return cvt_dotdiff_launch<float, float>(n, static_cast<const float*>(zin), static_cast<const float*>(r),
                                            static_cast<const float*>(zold), static_cast<float*>(znew), partial, st);
  if (dtype_in == F64)
    return cvt_dotdiff_launch<double, double>(n, static_cast<const double*>(zin), static_cast<const double*>(r),
                                              static_cast<const double*>(zold), static_cast<double*>(znew), partial, st);
  return cvt_dotdiff_launch<float, double>(n, static_cast<const float*>(zin), static_cast<const double*>(r),
                                           static_cast<const double*>(zold), static_cast<double*>(znew), partial, st);
}

int vh_jacobi_dot(int dtype, size_t n, const void* r, const void* dinv, void* z, double* partial, void* stream) {
  VH_REQUIRE(dtype_ok(dtype) && r && dinv && z && partial && n >= 1, "vh_jacobi_dot: bad arguments");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == F64)
    return jacobi_dot_launch<double>(n, static_cast<const double*>(r), static_cast<const double*>(dinv),
                                     static_cast<double*>(z), partial, st);
  return jacobi_dot_launch<float>(n, static_cast<const float*>(r), static_cast<const float*>(dinv), static_cast<float*>(z),
                                  partial, st);
}

int vh_xpby(int dtype, size_t n, const void* z, void* p, const double* scal, void* zcopy, void* stream) {
  VH_REQUIRE(dtype_ok(dtype) && z && p && scal && n >= 1, "vh_xpby: bad arguments");
  VH_REQUIRE(zcopy != p && zcopy != z, "vh_xpby: zcopy aliases an operand");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == F64)
    return xpby_launch<double>(n, static_cast<const double*>(z), static_cast<double*>(p), scal, static_cast<double*>(zcopy), st);
  return xpby_launch<float>(n, static_cast<const float*>(z), static_cast<float*>(p), scal, static_cast<float*>(zcopy), st);
}

int vh_axpby(int dtype, size_t n, double a, const void* x, const void* w, double b, void* y, void* stream) {
  VH_REQUIRE(dtype_ok(dtype) && x && y && n >= 1, "vh_axpby: bad arguments");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == F64)
    return axpby_launch<double>(n, a, static_cast<const double*>(x), static_cast<const double*>(w), b, static_cast<double*>(y), st);
  return axpby_launch<float>(n, float(a), static_cast<const float*>(x), static_cast<const float*>(w), float(b),
                             static_cast<float*>(y), st);
}

int vh_convert(int dtype_in, int dtype_out, size_t n, const void* x, void* y, double scale, const double* dscale,
               int inv_scale, int accumulate, void* stream) {
  VH_REQUIRE(dtype_ok(dtype_in) && (dtype_ok(dtype_out) || dtype_out == BF16) && x && y && n >= 1,
             "vh_convert: bad arguments");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype_out == BF16) {
    VH_REQUIRE(dtype_in == F32 && !accumulate, "vh_convert: bfloat16 output from float32, without accumulate");
    return convert_bf16_launch(n, static_cast<const float*>(x), static_cast<uint16_t*>(y), scale, dscale, inv_scale, st);
  }
  if (dtype_in == F64 && dtype_out == F32)
    return convert_launch<double, float>(n, static_cast<const double*>(x), static_cast<float*>(y), scale, dscale, inv_scale, accumulate, st);
  if (dtype_in == F32 && dtype_out == F64)
    return convert_launch<float, double>(n, static_cast<const float*>(x), static_cast<double*>(y), scale, dscale, inv_scale, accumulate, st);
  if (dtype_in == F64)
    return convert_launch<double, double>(n, static_cast<const double*>(x), static_cast<double*>(y), scale, dscale, inv_scale, accumulate, st);
  return convert_launch<float, float>(n, static_cast<const float*>(x), static_cast<float*>(y), scale, dscale, inv_scale, accumulate, st);
}

int vh_fill_splitmix(int dtype, size_t n, unsigned long long seed, void* y, void* stream) {
  VH_REQUIRE(dtype_ok(dtype) && y && n >= 1, "vh_fill_splitmix: bad arguments");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == F64) return fill_splitmix_launch<double>(n, seed, static_cast<double*>(y), st);
  return fill_splitmix_launch<float>(n, seed, static_cast<float*>(y), st);
}

int vh_cone_filter(int dtype, const void* x, const void* pre_div, const void* post_div, void* y, int nx, int ny, int nz,
                   double rmin, void* stream) {
  VH_REQUIRE(dtype_ok(dtype) && x && y && x != y && dims_ok(nx, ny, nz), "vh_cone_filter: bad arguments");
  VH_REQUIRE(rmin > 0 && rmin < 3.0, "vh_cone_filter: radius %g outside (0, 3)", rmin);
  const int R = int(std::ceil(rmin)) - 1;
  hipStream_t st = static_cast<hipStream_t>(stream);
  auto fill = [&](auto& c) {
    c.nx = nx; c.ny = ny; c.nz = nz; c.n_off = 0;
    for (int dx = -R; dx <= R; ++dx)
      for (int dy = -R; dy <= R; ++dy)
        for (int dz = -R; dz <= R; ++dz) {
          const double w = rmin - std::sqrt(double(dx * dx + dy * dy + dz * dz));
          if (w > 0) {
            c.dx[c.n_off] = dx; c.dy[c.n_off] = dy; c.dz[c.n_off] = dz;
            c.w[c.n_off] = w;
            ++c.n_off;
          }
        }
  };
  if (dtype == F64) {
    ConeArgs<double> c{};
    fill(c);
    return cone_launch<double>(static_cast<const double*>(x), static_cast<const double*>(pre_div),
                               static_cast<const double*>(post_div), static_cast<double*>(y), c, st);
  }
  ConeArgs<float> c{};
  fill(c);
  return cone_launch<float>(static_cast<const float*>(x), static_cast<const float*>(pre_div),
                            static_cast<const float*>(post_div), static_cast<float*>(y), c, st);
}

int vh_oc_probe(int dtype, size_t n, const void* x, const void* dc, const void* dv, const void* w, double lmid,
                double move, void* xnew, double* partial, int nparts, void* stream) {
  VH_REQUIRE(dtype_ok(dtype) && x && dc && partial && n >= 1 && nparts >= 1, "vh_oc_probe: bad arguments");
  VH_REQUIRE(lmid > 0, "vh_oc_probe: multiplier must be positive");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == F64)
    return oc_probe_launch<double>(n, static_cast<const double*>(x), static_cast<const double*>(dc),
                                   static_cast<const double*>(dv), static_cast<const double*>(w), lmid, move,
                                   static_cast<double*>(xnew), partial, nparts, st);
  return oc_probe_launch<float>(n, static_cast<const float*>(x), static_cast<const float*>(dc),
                                static_cast<const float*>(dv), static_cast<const float*>(w), float(lmid), float(move),
                                static_cast<float*>(xnew), partial, nparts, st);
}

int vh_oc_probe_multi(int dtype, size_t n, const void* x, const void* dc, const void* dv, const void* w,
                      const double* lmids, int count, double move, double* partial, int nparts, void* stream) {
  VH_REQUIRE(dtype_ok(dtype) && x && dc && partial && lmids && n >= 1 && nparts >= 1,
             "vh_oc_probe_multi: bad arguments");
  VH_REQUIRE(count >= 1 && count <= OC_MAX_PROBES, "vh_oc_probe_multi: %d multipliers (1 .. %d)", count,
             OC_MAX_PROBES);
  OcLambdas lm{};
  lm.count = count;
  for (int k = 0; k < count; ++k) {
    VH_REQUIRE(lmids[k] > 0, "vh_oc_probe_multi: multiplier %d must be positive", k);
    lm.v[k] = lmids[k];
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == F64)
    return oc_probe_multi_launch<double>(n, static_cast<const double*>(x), static_cast<const double*>(dc),
                                         static_cast<const double*>(dv), static_cast<const double*>(w), lm, move,
                                         partial, nparts, st);
  return oc_probe_multi_launch<float>(n, static_cast<const float*>(x), static_cast<const float*>(dc),
                                      static_cast<const float*>(dv), static_cast<const float*>(w), lm, float(move),
                                      partial, nparts, st);
}

int vh_heaviside(size_t n, const double* rho_bar, double beta, double eta, double penal, double E0, double Emin,
                 double* rho, double* drho, double* E, double* partial, int nparts, void* stream) {
  VH_REQUIRE(rho_bar && n >= 1 && (rho || drho || E || partial), "vh_heaviside: bad arguments");
  VH_REQUIRE(!partial || nparts >= 1, "vh_heaviside: nparts");
  VH_REQUIRE(beta >= 0 && eta > 0 && eta < 1, "vh_heaviside: beta %g, eta %g", beta, eta);
  const double tb = std::tanh(beta * eta);
  const double den = tb + std::tanh(beta * (1.0 - eta));
  const int grid = partial ? nparts : int(std::min<size_t>((n + 255) / 256, 4096));
  return heaviside_launch(n, rho_bar, beta, eta, tb, den, penal, E0, Emin, rho, drho, E, partial, grid,
                          static_cast<hipStream_t>(stream));
}

int vh_oc_design(size_t n, const double* rho, const double* dc, const double* dv, double lmid, double move,
                 double rho_lb, double* rho_new, void* stream) {
  VH_REQUIRE(rho && dc && dv && rho_new && n >= 1 && rho_new != rho, "vh_oc_design: bad arguments");
  VH_REQUIRE(lmid > 0 && move > 0, "vh_oc_design: multiplier and move limit must be positive");
  return oc_design_launch(n, rho, dc, dv, lmid, move, rho_lb, rho_new, static_cast<hipStream_t>(stream));
}

}  // extern "C"
