// Batched launches: many voxel problems (and load cases) of different grid sizes in one launch
// per pass.  Every batched operation reads a descriptor table built once on the host and copied
// to device memory (the same table serves every replay of a captured iteration), and maps its
// workgroups to problems through a prefix-sum table of per-problem workgroup counts (a binary
// search per workgroup; the problem index is then uniform across the workgroup).
//
//   product batch     one parity_xmarch_batch launch over the tiles of all problems
//   segments          per-problem ranges of a concatenated vector: segmented dot / |x|^2 and
//                     the Polak-Ribiere pass, one partial per workgroup
//   segmented finish  one workgroup per problem sums its partials and runs the scalar
//                     recurrence on its own 16-slot scalar block
//   transfer batch    restriction or prolongation, one workgroup per row block of any problem
//   dense batch       the coarsest solves (explicit inverses of different orders)
#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "common.h"
#include "blas.h"
#include "grid_ops.h"
#include "matvec.h"
#include "voxelhip.h"

namespace vh {

constexpr int SEG_BLOCK = 256;
constexpr int SEG_ITEMS = 8;
constexpr int SEG_CHUNK = SEG_BLOCK * SEG_ITEMS;      // elements per workgroup of a segmented pass

enum BatchKind : int { BK_PRODUCT = 1, BK_SEGMENTS = 2, BK_TRANSFER = 3, BK_DENSE = 4 };

struct BatchHandle {
  int kind = 0;
  int dtype = 0, v_dtype = 0;
  int n = 0, total = 0;             // problems; prefix total (tiles, blocks, lines, row blocks)
  int mode = 0, mask_kind = 0, tile = 0, combine = 0, restrict_op = 0, assign = 0;
  size_t lds = 0;                   // dynamic LDS of a transfer batch
  void* descs = nullptr;            // device descriptor table
  int* offs = nullptr;              // device prefix table (n + 1)
  long long* eoff = nullptr;        // device element prefix (segments)
};

__device__ __forceinline__ int find_segment(const int* __restrict__ offs, int n, int b) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (offs[mid] <= b) lo = mid; else hi = mid;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

// ------------------------------------------------------------------------------ segments
template <typename T, int KIND>
__global__ void __launch_bounds__(SEG_BLOCK) seg_reduce_kernel(const long long* __restrict__ eoff,
                                                               const int* __restrict__ boff, int n,
                                                               const T* __restrict__ x, const T* __restrict__ y,
                                                               double* __restrict__ partial) {
  __shared__ double red[SEG_BLOCK / 64];
  const int b = blockIdx.x;
  const int i = find_segment(boff, n, b);
  const long long lo = eoff[i] + (long long)(b - boff[i]) * SEG_CHUNK;
  const long long hi = min(lo + SEG_CHUNK, eoff[i + 1]);
  // each thread takes SEG_ITEMS adjacent elements (wide loads, all in flight together)
  const long long k0 = lo + (long long)threadIdx.x * SEG_ITEMS;
  double acc = 0.0;
  if (k0 + SEG_ITEMS <= hi) {
    T xv[SEG_ITEMS], yv[SEG_ITEMS];
#pragma unroll
    for (int t = 0; t < SEG_ITEMS; ++t) xv[t] = x[k0 + t];
    if (KIND != RED_NORM2) {
#pragma unroll
      for (int t = 0; t < SEG_ITEMS; ++t) yv[t] = y[k0 + t];
    }
#pragma unroll
    for (int t = 0; t < SEG_ITEMS; ++t)
      acc += KIND == RED_NORM2 ? double(xv[t]) * double(xv[t]) : double(xv[t]) * double(yv[t]);
  } else {
    for (long long k = k0; k < hi; ++k) {
      const double xv = double(x[k]);
      acc += KIND == RED_NORM2 ? xv * xv : xv * double(y[k]);
    }
  }
  const double s = block_sum<double, SEG_BLOCK>(acc, red);
  if (threadIdx.x == 0) partial[b] = s;
}

template <typename Ti, typename Tv>
__global__ void __launch_bounds__(SEG_BLOCK) seg_dotdiff_kernel(const long long* __restrict__ eoff,
                                                                const int* __restrict__ boff, int n, int total,
                                                                const Ti* __restrict__ zin, const Tv* __restrict__ r,
                                                                const Tv* __restrict__ zold, Tv* __restrict__ znew,
                                                                double* __restrict__ partial) {
  __shared__ double red[2][SEG_BLOCK / 64];
  const int b = blockIdx.x;
  const int i = find_segment(boff, n, b);
  const long long lo = eoff[i] + (long long)(b - boff[i]) * SEG_CHUNK;
  const long long hi = min(lo + SEG_CHUNK, eoff[i + 1]);
  // each thread takes SEG_ITEMS adjacent elements (wide loads and stores, all loads in flight)
  const long long k0 = lo + (long long)threadIdx.x * SEG_ITEMS;
  double s0 = 0.0, s1 = 0.0;
  if (k0 + SEG_ITEMS <= hi) {
    Tv zv[SEG_ITEMS], rv[SEG_ITEMS], ov[SEG_ITEMS];
#pragma unroll
    for (int t = 0; t < SEG_ITEMS; ++t) zv[t] = Tv(zin[k0 + t]);
#pragma unroll
    for (int t = 0; t < SEG_ITEMS; ++t) rv[t] = r[k0 + t];
#pragma unroll
    for (int t = 0; t < SEG_ITEMS; ++t) ov[t] = zold[k0 + t];
#pragma unroll
    for (int t = 0; t < SEG_ITEMS; ++t) znew[k0 + t] = zv[t];
#pragma unroll
    for (int t = 0; t < SEG_ITEMS; ++t) {
      const double z = double(zv[t]), rk = double(rv[t]);
      s0 += rk * z;
      s1 += rk * (z - double(ov[t]));
    }
  } else {
    for (long long k = k0; k < hi; ++k) {
      const Tv zk = Tv(zin[k]);
      const double z = double(zk), rk = double(r[k]);
      znew[k] = zk;
      s0 += rk * z;
      s1 += rk * (z - double(zold[k]));
    }
  }
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { red[0][w] = s0; red[1][w] = s1; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, c = 0.0;
    for (int k = 0; k < SEG_BLOCK / 64; ++k) { a += red[0][k]; c += red[1][k]; }
    partial[b] = a;
    partial[total + b] = c;
  }
}

// One workgroup per problem: partial layout [j][total], problem i owns [offs[i], offs[i+1]).
__global__ void __launch_bounds__(256) seg_finish_kernel(const double* __restrict__ partial, int total,
                                                         const int* __restrict__ offs, int count,
                                                         double* __restrict__ scal, int stride, int dst, int op) {
  __shared__ double red[2][4];
  const int i = blockIdx.x;
  const int lo = offs[i], hi = offs[i + 1];
  double acc[2] = {0.0, 0.0};
  for (int j = 0; j < count; ++j)
    for (int k = lo + threadIdx.x; k < hi; k += 256) acc[j] += partial[j * total + k];
  for (int j = 0; j < 2; ++j) acc[j] = wave_sum(acc[j]);
  if ((threadIdx.x & 63) == 0) { red[0][threadIdx.x >> 6] = acc[0]; red[1][threadIdx.x >> 6] = acc[1]; }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double v[2];
  for (int j = 0; j < 2; ++j) v[j] = red[j][0] + red[j][1] + red[j][2] + red[j][3];
  apply_scalar_op(scal + size_t(i) * stride, v, count, dst, op);
}

// ------------------------------------------------------------------------------ transfers
template <typename T>
struct TransferDesc {
  const T* src;
  T* dst;
  int nx, ny, nz;                   // fine element grid
  int fmask_kind, cmask_kind;
  const uint8_t* fmask;
  const uint8_t* cmask;
  int rows;                         // rows per block (restrict_rows / prolong_rows)
};

static_assert(TR_RESTRICT::NT == TR_PROLONG::NT, "one workgroup size for both transfer directions");

template <typename T, bool RESTRICT>
__global__ void __launch_bounds__(TR_RESTRICT::NT) transfer_batch_kernel(const TransferDesc<T>* __restrict__ d,
                                                                    const int* __restrict__ offs, int n, int assign) {
  extern __shared__ unsigned char smem_raw[];
  const int b = xcd_logical(blockIdx.x, gridDim.x);
  if (b >= offs[n]) return;
  const int i = find_segment(offs, n, b);
  const TransferDesc<T>& t = d[i];
  const int blk = b - offs[i];
  if (RESTRICT)
    restrict_block<T>(t.src, t.dst, t.nx, t.ny, t.nz, t.fmask_kind, t.fmask, t.cmask, t.cmask_kind, t.rows, blk,
                      reinterpret_cast<T*>(smem_raw));
  else
    prolong_block<T>(t.src, t.dst, t.nx, t.ny, t.nz, t.fmask_kind, t.fmask, assign, t.rows, blk,
                     reinterpret_cast<T*>(smem_raw));
}

// ------------------------------------------------------------------------------ dense solves
// One descriptor per group of up to DENSE_RHS right-hand sides that share the matrix (the load cases
// of one problem): the rows of A are read once for all of them.
constexpr int DENSE_RHS = 4;
template <typename TA, typename TV>
struct DenseDesc {
  const TA* A;
  const uint8_t* mask;
  int n, ld, k;                     // k right-hand sides (1 .. DENSE_RHS)
  const TV* r[DENSE_RHS];
  TV* y[DENSE_RHS];
};

// 4 rows per workgroup, one wave per row, as dense_mv_kernel (16-byte loads of A, several in flight);
// each right-hand side accumulates in the same order as a single-vector pass (chunk after chunk in
// column order, then the wave sum), so grouping does not change results.
template <typename TA, typename TV>
__global__ void __launch_bounds__(256) dense_batch_kernel(const DenseDesc<TA, TV>* __restrict__ d,
                                                          const int* __restrict__ offs, int nb) {
  constexpr int V = 16 / int(sizeof(TA));
  using Vec = typename std::conditional<sizeof(TA) == 8, double2, float4>::type;
  const int i = find_segment(offs, nb, blockIdx.x);
  const DenseDesc<TA, TV>& q = d[i];
  const int row = (blockIdx.x - offs[i]) * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= q.n) return;
  const TA* a = q.A + size_t(row) * q.ld;
  const int nfull = (q.n / V) * V;
  const int k = q.k;
  TA acc[DENSE_RHS];
#pragma unroll
  for (int c = 0; c < DENSE_RHS; ++c) acc[c] = TA(0);
  // the loads of UN chunks of the row in flight before they are accumulated (each right-hand
  // side's sum keeps the column order of one chunk after the other)
  constexpr int UN = 4;
  int j = lane * V;
  for (; j + (UN - 1) * 64 * V < nfull; j += UN * 64 * V) {
    Vec x[UN];
#pragma unroll
    for (int u = 0; u < UN; ++u) x[u] = *reinterpret_cast<const Vec*>(a + j + u * 64 * V);
#pragma unroll
    for (int u = 0; u < UN; ++u) {
      const TA* e = reinterpret_cast<const TA*>(&x[u]);
#pragma unroll
      for (int c = 0; c < DENSE_RHS; ++c) {
        if (c < k) {
          const TV* r = q.r[c] + j + u * 64 * V;
#pragma unroll
          for (int t = 0; t < V; ++t) acc[c] += e[t] * TA(r[t]);
        }
      }
    }
  }
  for (; j < nfull; j += 64 * V) {
    const Vec x = *reinterpret_cast<const Vec*>(a + j);
    const TA* e = reinterpret_cast<const TA*>(&x);
#pragma unroll
    for (int c = 0; c < DENSE_RHS; ++c) {
      if (c < k) {
        const TV* r = q.r[c];
#pragma unroll
        for (int t = 0; t < V; ++t) acc[c] += e[t] * TA(r[j + t]);
      }
    }
  }
  for (int t = nfull + lane; t < q.n; t += 64) {
#pragma unroll
    for (int c = 0; c < DENSE_RHS; ++c)
      if (c < k) acc[c] += a[t] * TA(q.r[c][t]);
  }
  const bool fixed = q.mask && !q.mask[row];
#pragma unroll
  for (int c = 0; c < DENSE_RHS; ++c) {
    if (c < k) {
      const TA s = wave_sum(acc[c]);
      if (lane == 0) q.y[c][row] = fixed ? TV(0) : TV(s);
    }
  }
}

static int upload(void** dst, const void* src, size_t bytes) {
  if (hipMalloc(dst, bytes) != hipSuccess) return fail(-1, "batch: hipMalloc of %zu bytes failed", bytes);
  return hip_status(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice), "batch: descriptor upload");
}

static int drop(BatchHandle* h, int r) {
  hipFree(h->descs);
  delete h;
  return r;
}

static int finish_handle(BatchHandle* h, const std::vector<int>& prefix, void** out) {
  h->n = int(prefix.size()) - 1;
  h->total = prefix.back();
  if (int r = upload(reinterpret_cast<void**>(&h->offs), prefix.data(), prefix.size() * sizeof(int))) {
    hipFree(h->descs);
    delete h;
    return r;
  }
  *out = h;
  return 0;
}

template <typename T>
static int product_create(int mask_kind, int tile, int n, const vh_product_item* items, BatchHandle* h,
                          int* tiles_out) {
  std::vector<MatvecParams<T>> descs(n);
  std::vector<int> prefix(n + 1, 0);
  // x chunk lengths planned for the batch as a whole (items that give their own lx keep it)
  std::vector<int> bnx(n), bny(n), bnz(n), blx(n);
  for (int i = 0; i < n; ++i) {
    bnx[i] = items[i].nx;
    bny[i] = items[i].ny;
    bnz[i] = items[i].nz;
  }
  matvec_batch_chunks<T>(n, bnx.data(), bny.data(), bnz.data(), tile, blx.data());
  for (int i = 0; i < n; ++i) {
    const vh_product_item& it = items[i];
    MatvecParams<T>& p = descs[i];
    p = MatvecParams<T>{};
    double k8[8];
    if (!cube_pattern(it.blocks72, k8)) return fail(-2, "vh_product_batch_create: item %d is not a voxel cube", i);
    p.cube = 1;
    for (int k = 0; k < 8; ++k) p.coef[k] = T(k8[k]);
    p.u = it.u;
    p.E = static_cast<const T*>(it.E);
    p.y = static_cast<T*>(it.y);
    p.mask = it.mask;
    p.nx = it.nx; p.ny = it.ny; p.nz = it.nz;
    const vh_epilogue& ep = it.ep;
    p.b = static_cast<const T*>(ep.b);
    p.dinv = static_cast<const T*>(ep.dinv);
    p.d = static_cast<T*>(ep.d);
    p.omega = T(ep.omega);
    p.cheb_a = T(ep.cheb_a);
    p.cheb_c = T(ep.cheb_c);
    for (int c = 0; c < 3; ++c) p.kdi[c] = ep.kdiag[c] > 0 ? T(1.0 / ep.kdiag[c]) : T(0);
    p.flags = ep.flags;
    p.w = ep.w;
    p.sdev = ep.sdev;
    p.sidx = ep.sidx;
    p.u_out = static_cast<T*>(ep.u_out);
    p.y32 = static_cast<float*>(ep.y32);
    p.fixed_zero = ep.fixed_zero;
    p.load_scale = T(ep.load_scale);
    p.prev = static_cast<const T*>(ep.prev);
    const MatvecPlan pl = matvec_batch_plan<T>(it.nx, it.ny, it.nz, ep.lx > 0 ? ep.lx : blx[i], tile);
    p.tile = pl.tile;
    p.lx = pl.lx;
    p.tiles_y = pl.tiles_y;
    p.tiles_z = pl.tiles_z;
    p.tile_ty = pl.tile_ty;
    p.chunks_x = pl.chunks_x;
    p.n_tiles = pl.n_tiles;
    prefix[i + 1] = prefix[i] + pl.n_tiles;
    // all items share the partial array: item i writes tiles [prefix[i], prefix[i+1])
    p.partial = ep.partial ? ep.partial + prefix[i] : nullptr;
    if (tiles_out) tiles_out[i] = pl.n_tiles;
  }
  h->tile = (tile == MV_TILE_8x32 || tile == MV_TILE_16x16) ? tile : MV_TILE_16x32;
  return upload(&h->descs, descs.data(), descs.size() * sizeof(MatvecParams<T>));
}

template <typename T>
static int transfer_create(int n, const vh_transfer_item* items, BatchHandle* h, std::vector<int>& prefix) {
  std::vector<TransferDesc<T>> descs(n);
  size_t lds = 0;
  for (int i = 0; i < n; ++i) {
    const vh_transfer_item& it = items[i];
    // the problems of a batch fill the chip together: rows grow with no block-count limit
    const int rows = h->restrict_op ? restrict_rows<TR_RESTRICT>(it.nx, it.ny, it.nz, 0)
                                    : prolong_rows<TR_PROLONG>(it.nx, it.ny, it.nz, 0);
    descs[i] = TransferDesc<T>{static_cast<const T*>(it.src), static_cast<T*>(it.dst), it.nx, it.ny, it.nz,
                               it.fmask_kind, it.cmask_kind, it.fmask, it.cmask, rows};
    const TransferGeom g = h->restrict_op ? restrict_geom<TR_RESTRICT>(it.nx, it.ny, it.nz, rows)
                                          : prolong_geom<TR_PROLONG>(it.nx, it.ny, it.nz, rows);
    prefix[i + 1] = prefix[i] + g.nblocks;
    lds = std::max(lds, g.lds);
  }
  h->lds = lds * sizeof(T);
  return upload(&h->descs, descs.data(), descs.size() * sizeof(TransferDesc<T>));
}

template <typename TA, typename TV>
static int dense_create(int n, const vh_dense_item* items, BatchHandle* h, std::vector<int>& prefix) {
  // consecutive items with the same matrix (pointer, size, stride and mask) share a descriptor
  std::vector<DenseDesc<TA, TV>> descs;
  prefix.assign(1, 0);
  for (int i = 0; i < n; ++i) {
    const vh_dense_item& it = items[i];
    const TA* A = static_cast<const TA*>(it.A);
    if (descs.empty() || descs.back().A != A || descs.back().n != it.n || descs.back().ld != it.ld ||
        descs.back().mask != it.mask || descs.back().k == DENSE_RHS) {
      DenseDesc<TA, TV> q{};
      q.A = A;
      q.mask = it.mask;
      q.n = it.n;
      q.ld = it.ld;
      descs.push_back(q);
      prefix.push_back(prefix.back() + (it.n + 3) / 4);
    }
    DenseDesc<TA, TV>& q = descs.back();
    q.r[q.k] = static_cast<const TV*>(it.r);
    q.y[q.k] = static_cast<TV*>(it.y);
    ++q.k;
  }
  return upload(&h->descs, descs.data(), descs.size() * sizeof(DenseDesc<TA, TV>));
}

}  // namespace vh

using namespace vh;

#define VH_REQUIRE(cond, ...) \
  do {                        \
    if (!(cond)) return fail(-2, __VA_ARGS__); \
  } while (0)

static bool dt_ok(int d) { return d == F64 || d == F32; }

extern "C" {

int vh_product_batch_create(int dtype, int mask_kind, int tile, int n, const vh_product_item* items, void** handle,
                            int* tiles_out) {
  VH_REQUIRE(dt_ok(dtype) && n >= 1 && items && handle, "vh_product_batch_create: bad arguments");
  VH_REQUIRE(mask_kind == MASK_X0 || mask_kind == MASK_ARRAY, "vh_product_batch_create: mask kind %d", mask_kind);
  const int mode = items[0].ep.mode;
  // load-time combination: 0 none, 1 w of the compute type, 2 an FP32 w in an FP64 RESID pass
  // (the mixed-precision Krylov loop's search direction, flag MV_W_F32)
  const int combine = items[0].ep.w == nullptr ? 0 : ((items[0].ep.flags & MV_W_F32) ? 2 : 1);
  VH_REQUIRE(combine != 2 || (dtype == F64 && mode == MODE_RESID),
             "vh_product_batch_create: an FP32 w needs the FP64 RESID pass");
  VH_REQUIRE(mode == MODE_RAW || mode == MODE_RESID || mode == MODE_DOT || mode == MODE_JACOBI || mode == MODE_CHEB3,
             "vh_product_batch_create: mode %d", mode);
  VH_REQUIRE(!combine || mode == MODE_RESID || mode == MODE_DOT || mode == MODE_CHEB3,
             "vh_product_batch_create: combination needs DOT/RESID/CHEB3");
  for (int i = 0; i < n; ++i) {
    const vh_product_item& it = items[i];
    VH_REQUIRE(it.blocks72 && it.E && it.u && (it.y || mode == MODE_DOT || (mode == MODE_RESID && it.ep.y32)) &&
                   it.u != it.y,
               "vh_product_batch_create: item %d pointers", i);
    VH_REQUIRE(it.nx >= 1 && it.ny >= 1 && it.nz >= 1 && it.nz <= 2048 &&
                   3.0 * (it.nx + 1.0) * (it.ny + 1.0) * (it.nz + 1.0) < 2147483648.0,
               "vh_product_batch_create: item %d grid", i);
    VH_REQUIRE(it.ep.mode == mode && (it.ep.w == nullptr ? 0 : ((it.ep.flags & MV_W_F32) ? 2 : 1)) == combine,
               "vh_product_batch_create: item %d differs", i);
    VH_REQUIRE((it.ep.flags & ~(MV_PREV_ZERO | MV_W_F32)) == 0, "vh_product_batch_create: item %d flags", i);
    VH_REQUIRE(mask_kind != MASK_ARRAY || it.mask, "vh_product_batch_create: item %d mask", i);
    VH_REQUIRE(!it.ep.prev || (mode == MODE_CHEB3 && !combine && it.ep.prev != it.y),
               "vh_product_batch_create: item %d prev", i);
    VH_REQUIRE(!(mode == MODE_RESID || mode == MODE_JACOBI || mode == MODE_CHEB3) || it.ep.b,
               "vh_product_batch_create: item %d b", i);
    VH_REQUIRE(mode != MODE_DOT || it.ep.partial, "vh_product_batch_create: item %d partial", i);
    VH_REQUIRE(!combine || ((it.ep.sdev || mode == MODE_CHEB3) && it.ep.u_out && it.ep.u_out != it.u &&
                            it.ep.u_out != it.y && it.ep.w != it.y && it.ep.u_out != it.ep.w && it.ep.u_out != it.ep.b),
               "vh_product_batch_create: item %d combination operands", i);
    VH_REQUIRE(!combine || mode != MODE_CHEB3 || (it.ep.flags & MV_PREV_ZERO) || it.ep.cheb_c == 0.0,
               "vh_product_batch_create: item %d: a CHEB3 pass with the load-time product starts from zero", i);
    VH_REQUIRE(!it.ep.y32 || mode == MODE_RESID, "vh_product_batch_create: item %d y32", i);
    if ((mode == MODE_JACOBI || mode == MODE_CHEB3) && !it.ep.dinv)
      VH_REQUIRE(it.ep.kdiag[0] > 0 && it.ep.kdiag[1] > 0 && it.ep.kdiag[2] > 0,
                 "vh_product_batch_create: item %d kdiag", i);
  }
  BatchHandle* h = new BatchHandle();
  h->kind = BK_PRODUCT;
  h->dtype = dtype;
  h->mode = mode;
  h->mask_kind = mask_kind;
  h->combine = combine;
  std::vector<int> prefix(n + 1, 0);
  // per-problem plans (prefix filled by the typed builder through tiles_out)
  std::vector<int> tiles(n, 0);
  const int r = dtype == F64 ? product_create<double>(mask_kind, tile, n, items, h, tiles.data())
                             : product_create<float>(mask_kind, tile, n, items, h, tiles.data());
  if (r) return drop(h, r);
  for (int i = 0; i < n; ++i) prefix[i + 1] = prefix[i] + tiles[i];
  if (tiles_out) std::memcpy(tiles_out, tiles.data(), n * sizeof(int));
  return finish_handle(h, prefix, handle);
}

int vh_segments_create(int n, const long long* lengths, void** handle) {
  VH_REQUIRE(n >= 1 && lengths && handle, "vh_segments_create: bad arguments");
  std::vector<long long> eoff(n + 1, 0);
  std::vector<int> boff(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    VH_REQUIRE(lengths[i] >= 1, "vh_segments_create: segment %d is empty", i);
    eoff[i + 1] = eoff[i] + lengths[i];
    boff[i + 1] = boff[i] + int((lengths[i] + SEG_CHUNK - 1) / SEG_CHUNK);
  }
  BatchHandle* h = new BatchHandle();
  h->kind = BK_SEGMENTS;
  const int r = upload(reinterpret_cast<void**>(&h->eoff), eoff.data(), eoff.size() * sizeof(long long));
  h->descs = h->eoff;                    // released with the handle
  if (r) return drop(h, r);
  return finish_handle(h, boff, handle);
}

int vh_transfer_batch_create(int dtype, int restrict_op, int assign, int n, const vh_transfer_item* items,
                             void** handle) {
  VH_REQUIRE(dt_ok(dtype) && n >= 1 && items && handle, "vh_transfer_batch_create: bad arguments");
  for (int i = 0; i < n; ++i) {
    const vh_transfer_item& it = items[i];
    VH_REQUIRE(it.src && it.dst && it.src != it.dst, "vh_transfer_batch_create: item %d pointers", i);
    VH_REQUIRE(it.nx >= 2 && it.ny >= 2 && it.nz >= 2 && it.nx % 2 == 0 && it.ny % 2 == 0 && it.nz % 2 == 0 &&
                   it.nz <= 2048,
               "vh_transfer_batch_create: item %d fine grid %d x %d x %d", i, it.nx, it.ny, it.nz);
    VH_REQUIRE((it.fmask_kind != MASK_ARRAY || it.fmask) && (it.cmask_kind != MASK_ARRAY || it.cmask),
               "vh_transfer_batch_create: item %d mask", i);
  }
  BatchHandle* h = new BatchHandle();
  h->kind = BK_TRANSFER;
  h->dtype = dtype;
  h->restrict_op = restrict_op != 0;
  h->assign = assign != 0;
  std::vector<int> prefix(n + 1, 0);
  const int r = dtype == F64 ? transfer_create<double>(n, items, h, prefix) : transfer_create<float>(n, items, h, prefix);
  if (r) return drop(h, r);
  return finish_handle(h, prefix, handle);
}

int vh_dense_batch_create(int a_dtype, int v_dtype, int n, const vh_dense_item* items, void** handle) {
  VH_REQUIRE(dt_ok(a_dtype) && dt_ok(v_dtype) && !(a_dtype == F32 && v_dtype == F64) && n >= 1 && items && handle,
             "vh_dense_batch_create: bad arguments");
  for (int i = 0; i < n; ++i) {
    const vh_dense_item& it = items[i];
    VH_REQUIRE(it.A && it.r && it.y && it.r != it.y && it.n >= 1 && it.ld >= it.n && it.ld % 4 == 0 &&
                   (reinterpret_cast<uintptr_t>(it.A) & 15) == 0,
               "vh_dense_batch_create: item %d", i);
  }
  BatchHandle* h = new BatchHandle();
  h->kind = BK_DENSE;
  h->dtype = a_dtype;
  h->v_dtype = v_dtype;
  std::vector<int> prefix(n + 1, 0);
  int r;
  if (a_dtype == F64 && v_dtype == F64) r = dense_create<double, double>(n, items, h, prefix);
  else if (a_dtype == F64) r = dense_create<double, float>(n, items, h, prefix);
  else r = dense_create<float, float>(n, items, h, prefix);
  if (r) return drop(h, r);
  return finish_handle(h, prefix, handle);
}

int vh_batch_launch(void* handle, void* stream) {
  VH_REQUIRE(handle, "vh_batch_launch: null handle");
  const BatchHandle* h = static_cast<const BatchHandle*>(handle);
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (h->kind) {
    case BK_PRODUCT: {
      MatvecBatch b{h->mode, h->mask_kind, h->tile, h->combine, h->n, h->total, h->descs, h->offs};
      return h->dtype == F64 ? matvec_batch_launch<double>(b, st) : matvec_batch_launch<float>(b, st);
    }
    case BK_TRANSFER: {
      const dim3 g(unsigned((h->total + 7) / 8) * 8), blk(TR_RESTRICT::NT);
      if (h->dtype == F64) {
        const auto* d = static_cast<const TransferDesc<double>*>(h->descs);
        if (h->restrict_op) hipLaunchKernelGGL((transfer_batch_kernel<double, true>), g, blk, h->lds, st, d, h->offs, h->n, h->assign);
        else hipLaunchKernelGGL((transfer_batch_kernel<double, false>), g, blk, h->lds, st, d, h->offs, h->n, h->assign);
      } else {
        const auto* d = static_cast<const TransferDesc<float>*>(h->descs);
        if (h->restrict_op) hipLaunchKernelGGL((transfer_batch_kernel<float, true>), g, blk, h->lds, st, d, h->offs, h->n, h->assign);
        else hipLaunchKernelGGL((transfer_batch_kernel<float, false>), g, blk, h->lds, st, d, h->offs, h->n, h->assign);
      }
      return hip_status(hipGetLastError(), "transfer_batch_kernel");
    }
    case BK_DENSE: {
      const dim3 g(h->total), blk(256);
      if (h->dtype == F64 && h->v_dtype == F64)
        hipLaunchKernelGGL((dense_batch_kernel<double, double>), g, blk, 0, st,
                           static_cast<const DenseDesc<double, double>*>(h->descs), h->offs, h->n);
      else if (h->dtype == F64)
        hipLaunchKernelGGL((dense_batch_kernel<double, float>), g, blk, 0, st,
                           static_cast<const DenseDesc<double, float>*>(h->descs), h->offs, h->n);
      else
        hipLaunchKernelGGL((dense_batch_kernel<float, float>), g, blk, 0, st,
                           static_cast<const DenseDesc<float, float>*>(h->descs), h->offs, h->n);
      return hip_status(hipGetLastError(), "dense_batch_kernel");
    }
    default:
      return fail(-2, "vh_batch_launch: handle kind %d has no single launch", h->kind);
  }
}

int vh_batch_offsets(void* handle, const int** offs, int* n, int* total) {
  VH_REQUIRE(handle, "vh_batch_offsets: null handle");
  const BatchHandle* h = static_cast<const BatchHandle*>(handle);
  if (offs) *offs = h->offs;
  if (n) *n = h->n;
  if (total) *total = h->total;
  return 0;
}

int vh_batch_destroy(void* handle) {
  if (!handle) return 0;
  BatchHandle* h = static_cast<BatchHandle*>(handle);
  hipFree(h->descs);
  hipFree(h->offs);
  delete h;
  return 0;
}

int vh_seg_reduce(void* segments, int dtype, int kind, const void* x, const void* y, double* partial, void* stream) {
  VH_REQUIRE(segments && x && partial && dt_ok(dtype), "vh_seg_reduce: bad arguments");
  const BatchHandle* h = static_cast<const BatchHandle*>(segments);
  VH_REQUIRE(h->kind == BK_SEGMENTS, "vh_seg_reduce: not a segment table");
  VH_REQUIRE(kind == RED_NORM2 || (kind == RED_DOT && y), "vh_seg_reduce: kind %d", kind);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 g(h->total), blk(SEG_BLOCK);
  if (dtype == F64) {
    const double* xd = static_cast<const double*>(x);
    const double* yd = static_cast<const double*>(y);
    if (kind == RED_NORM2) hipLaunchKernelGGL((seg_reduce_kernel<double, RED_NORM2>), g, blk, 0, st, h->eoff, h->offs, h->n, xd, yd, partial);
    else hipLaunchKernelGGL((seg_reduce_kernel<double, RED_DOT>), g, blk, 0, st, h->eoff, h->offs, h->n, xd, yd, partial);
  } else {
    const float* xf = static_cast<const float*>(x);
    const float* yf = static_cast<const float*>(y);
    if (kind == RED_NORM2) hipLaunchKernelGGL((seg_reduce_kernel<float, RED_NORM2>), g, blk, 0, st, h->eoff, h->offs, h->n, xf, yf, partial);
    else hipLaunchKernelGGL((seg_reduce_kernel<float, RED_DOT>), g, blk, 0, st, h->eoff, h->offs, h->n, xf, yf, partial);
  }
  return hip_status(hipGetLastError(), "seg_reduce_kernel");
}

int vh_seg_dotdiff(void* segments, int dtype_in, int dtype_v, const void* zin, const void* r, const void* zold,
                   void* znew, double* partial, void* stream) {
  VH_REQUIRE(segments && zin && r && zold && znew && partial && dt_ok(dtype_in) && dt_ok(dtype_v),
             "vh_seg_dotdiff: bad arguments");
  VH_REQUIRE(znew != zin && znew != zold && znew != r, "vh_seg_dotdiff: znew aliases an input");
  VH_REQUIRE(!(dtype_in == F64 && dtype_v == F32), "vh_seg_dotdiff: FP64 input into FP32 vectors");
  const BatchHandle* h = static_cast<const BatchHandle*>(segments);
  VH_REQUIRE(h->kind == BK_SEGMENTS, "vh_seg_dotdiff: not a segment table");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 g(h->total), blk(SEG_BLOCK);
  if (dtype_v == F32)
    hipLaunchKernelGGL((seg_dotdiff_kernel<float, float>), g, blk, 0, st, h->eoff, h->offs, h->n, h->total,
                       static_cast<const float*>(zin), static_cast<const float*>(r), static_cast<const float*>(zold),
                       static_cast<float*>(znew), partial);
  else if (dtype_in == F32)
    hipLaunchKernelGGL((seg_dotdiff_kernel<float, double>), g, blk, 0, st, h->eoff, h->offs, h->n, h->total,
                       static_cast<const float*>(zin), static_cast<const double*>(r), static_cast<const double*>(zold),
                       static_cast<double*>(znew), partial);
  else
    hipLaunchKernelGGL((seg_dotdiff_kernel<double, double>), g, blk, 0, st, h->eoff, h->offs, h->n, h->total,
                       static_cast<const double*>(zin), static_cast<const double*>(r), static_cast<const double*>(zold),
                       static_cast<double*>(znew), partial);
  return hip_status(hipGetLastError(), "seg_dotdiff_kernel");
}

int vh_seg_finish(const double* partial, int total, const int* offs, int n, int count, double* scal, int stride,
                  int dst, int op, void* stream) {
  VH_REQUIRE(partial && offs && scal && n >= 1 && total >= 1 && (count == 1 || count == 2) && stride >= 16 &&
                 dst >= 0 && dst + count <= stride && op >= OP_NONE && op <= OP_RESNORM,
             "vh_seg_finish: bad arguments");
  hipLaunchKernelGGL(seg_finish_kernel, dim3(n), dim3(256), 0, static_cast<hipStream_t>(stream), partial, total, offs,
                     count, scal, stride, dst, op);
  return hip_status(hipGetLastError(), "seg_finish_kernel");
}

}  // extern "C"
