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
//  * Node planes are staged through LDS with a one-plane register prefetch: lane k loads node k
//    of the staged plane (rows of (x, y, z) triples, contiguous across a wave) with one wide load
//    and stores its three values into three per-component arrays, so the stores and the reads
//    of neighbouring nodes are unit-stride in LDS.  Each element column stores its four corner
//    contributions in LDS slots numbered by its thread, and the owner of a node gathers them
//    from its four incident columns in a fixed order (unit-stride writes and reads across a
//    wave; results are bitwise reproducible).
//  * The epilogue fuses the constrained operator (M K M + I - M), the residual, a damped
//    Jacobi sweep, a Chebyshev step or a p.Ap partial sum into the single pass over y.
//  * Workgroup ids are remapped so that each XCD (8 per MI355X, own L2) receives a contiguous
//    range of tiles with the x chunk slowest: halo rows, columns and planes shared by
//    neighbouring tiles are then mostly read from the same L2.
//
// This file is compiled twice, through matvec_f64.hip (VH_MATVEC_F64: the FP64 instantiations)
// and matvec_f32.hip (VH_MATVEC_F32: FP32 with FP32 or BF16 storage), so the two halves build
// in parallel.
#include <algorithm>
#include <type_traits>
#include <vector>

#include "common.h"
#include "matvec.h"

namespace vh {

// Block product of one element: f = E C v on the modal coefficients v[mode][comp].
// General bricks: the 8 blocks of 3x3 (72 coefficients).  Voxel cubes (any edge length):
// the blocks have the fixed pattern checked on the host (capi.hip, cube_pattern) with 8
// distinct values, p, q: B0 = (p-q) I + q 11^T;  a, b: B1, B2, B4 (one zero row/column each,
// the rigid translations);  c, d: B3, B5, B6 (a rank-one 2x2 part and one diagonal entry);
// e, f: B7 = (e-f) I + f 11^T.  E is folded into the 8 values once per element.
template <typename T, bool CUBE>
__device__ __forceinline__ void element_blocks(const T (&v)[8][3], T Ee, const MatvecParams<T>& a, T (&f)[8][3]) {
  if constexpr (CUBE) {
    const T* k = a.coef;              // [p-q, q, a, b, c, d, e-f, f]
    const T Epq = Ee * k[0], Eq = Ee * k[1], Ea = Ee * k[2], Eb = Ee * k[3];
    const T Ec = Ee * k[4], Ed = Ee * k[5], Eef = Ee * k[6], Ef = Ee * k[7];
    {  // g = 0: (v[1][0], v[2][1], v[4][2])
      const T s = (v[1][0] + v[2][1]) + v[4][2];
      const T qs = Eq * s;
      f[1][0] = Epq * v[1][0] + qs;
      f[2][1] = Epq * v[2][1] + qs;
      f[4][2] = Epq * v[4][2] + qs;
    }
    {  // g = 7: (v[6][0], v[5][1], v[3][2])
      const T s = (v[6][0] + v[5][1]) + v[3][2];
      const T fs = Ef * s;
      f[6][0] = Eef * v[6][0] + fs;
      f[5][1] = Eef * v[5][1] + fs;
      f[3][2] = Eef * v[3][2] + fs;
    }
    // g = 1: (v[0][0] unused, v[3][1], v[5][2])
    f[0][0] = T(0);
    f[3][1] = Ea * v[3][1] + Eb * v[5][2];
    f[5][2] = Eb * v[3][1] + Ea * v[5][2];
    // g = 2: (v[3][0], v[0][1] unused, v[6][2])
    f[3][0] = Ea * v[3][0] + Eb * v[6][2];
    f[0][1] = T(0);
    f[6][2] = Eb * v[3][0] + Ea * v[6][2];
    // g = 4: (v[5][0], v[6][1], v[0][2] unused)
    f[5][0] = Ea * v[5][0] + Eb * v[6][1];
    f[6][1] = Eb * v[5][0] + Ea * v[6][1];
    f[0][2] = T(0);
    {  // g = 3: (v[2][0], v[1][1], v[7][2])
      const T t = Ec * (v[2][0] + v[1][1]);
      f[2][0] = t;
      f[1][1] = t;
      f[7][2] = Ed * v[7][2];
    }
    {  // g = 5: (v[4][0], v[7][1], v[1][2])
      const T t = Ec * (v[4][0] + v[1][2]);
      f[4][0] = t;
      f[1][2] = t;
      f[7][1] = Ed * v[7][1];
    }
    {  // g = 6: (v[7][0], v[4][1], v[2][2])
      const T t = Ec * (v[4][1] + v[2][2]);
      f[4][1] = t;
      f[2][2] = t;
      f[7][0] = Ed * v[7][0];
    }
  } else {
#pragma unroll
    for (int g = 0; g < 8; ++g) {
      const T a0 = Ee * v[g ^ 1][0], a1 = Ee * v[g ^ 2][1], a2 = Ee * v[g ^ 4][2];
      f[g ^ 1][0] = a.coef[9 * g + 0] * a0 + a.coef[9 * g + 1] * a1 + a.coef[9 * g + 2] * a2;
      f[g ^ 2][1] = a.coef[9 * g + 3] * a0 + a.coef[9 * g + 4] * a1 + a.coef[9 * g + 5] * a2;
      f[g ^ 4][2] = a.coef[9 * g + 6] * a0 + a.coef[9 * g + 7] * a1 + a.coef[9 * g + 8] * a2;
    }
  }
}

// 1 / x: IEEE division in FP64; the hardware reciprocal (1 ulp) in FP32, where it only scales
// a smoothing correction.
__device__ __forceinline__ double reciprocal(double x) { return 1.0 / x; }
__device__ __forceinline__ float reciprocal(float x) { return __builtin_amdgcn_rcpf(x); }

// One workgroup's share of the product: tile `logical` (y tile fastest, then z tile, then x
// chunk) of the grid described by `a`.
template <typename T, typename S, int MODE, int MASK, typename TL, bool CUBE, int LOADOP>
__device__ __forceinline__ void xmarch_tile(const MatvecParams<T>& a, const int logical) {
  constexpr int NZ = TL::NZ, EZ = TL::EZ, BLOCK = TL::BLOCK, PLANE = TL::PLANE;
  // staged node planes ix & 1 (input), one array per component (node-major within it): reads
  // of neighbouring nodes by neighbouring lanes are unit-stride (FP64 values at the 24-byte
  // stride of the HBM layout put two lanes of a wave on one LDS bank)
  __shared__ T su[2][3][PLANE];
  // corner contributions and column modulus sums are indexed by the writing thread (unit stride
  // for the writes and for the owner's reads), plus one zero pad slot past every writer; corner
  // (1, 1) of a column is the node its own thread owns and stays in registers
  constexpr int SCN = BLOCK + EZ + 2;
  __shared__ T sc[3][3][SCN];         // corners (0,0), (0,1), (1,0) of each element column
  __shared__ T se[SCN];               // per element column: E(ex-1) + E(ex), for D^-1 on the fly
  __shared__ T sE[BLOCK];             // this thread's prefetched modulus, parked across the barrier
  __shared__ double red[BLOCK / 64];
  constexpr bool NEEDS_DINV = MODE == MODE_JACOBI || MODE == MODE_CHEB || MODE == MODE_CHEB3 || MODE == MODE_CHEB3D;
  // FP64 smoothing and residual passes, and FP64 passes that combine two vectors at load time,
  // recompute the left face from the still-staged plane instead of carrying it in 24 registers,
  // which keeps them within the 168 VGPRs of 3 waves per SIMD
  constexpr bool RECOMPUTE_FL = sizeof(T) == 8 && (NEEDS_DINV || MODE == MODE_RESID || LOADOP != LOAD_NONE);
  const bool dinv_from_E = NEEDS_DINV && a.dinv == nullptr;

  const int tz = logical % a.tiles_z;
  const int ty = (logical / a.tiles_z) % a.tiles_y;
  const int cx = logical / (a.tiles_z * a.tiles_y);

  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int ny1 = ny + 1, nz1 = nz + 1;
  const int tye = a.tile_ty;                       // owned node rows of this launch's y tiles
  const int oy0 = ty * tye, oy1 = (ty == a.tiles_y - 1) ? ny1 : oy0 + tye;
  const int oz0 = tz * TL::TZ, oz1 = (tz == a.tiles_z - 1) ? nz1 : oz0 + TL::TZ;
  const int ox0 = cx * a.lx, ox1 = (cx == a.chunks_x - 1) ? nx + 1 : ox0 + a.lx;
  const int ybase = oy0 - 1, zbase = oz0 - 1;      // local node j <-> global base + j

  const int tid = threadIdx.x;
  const int ly = tid / EZ, lz = tid - (tid / EZ) * EZ;
  const int ey = ybase + ly, ez = zbase + lz;
  const bool col_in_tile = tid < TL::NCOL && ly <= tye;
  const bool col_ok = col_in_tile && ey >= 0 && ey < ny && ez >= 0 && ez < nz;
  const int iy_own = oy0 + ly, iz_own = oz0 + lz;  // node (ly+1, lz+1) in local coordinates
  const bool node_own = col_in_tile && iy_own < oy1 && iz_own < oz1;

  // zero the contribution slots once: slots without a writer column must read as zero
  {
    T* flat = &sc[0][0][0];
    for (int i = tid; i < 9 * SCN; i += BLOCK) flat[i] = T(0);
    if (dinv_from_E)
      for (int i = tid; i < SCN; i += BLOCK) se[i] = T(0);
  }

  const S* __restrict__ u = reinterpret_cast<const S*>(a.u);
  const T* __restrict__ wv = reinterpret_cast<const T*>(a.w);     // combination operand, compute type
  const float* __restrict__ wf = reinterpret_cast<const float*>(a.w);   // (AXPY32: stored in FP32)
  const T* __restrict__ E = a.E;
  constexpr bool COMBINE = LOADOP != LOAD_NONE;
  constexpr bool AXPY = LOADOP == LOAD_AXPY || LOADOP == LOAD_AXPY32;
  auto w_at = [&](size_t dof) -> T { return LOADOP == LOAD_AXPY32 ? T(wf[dof]) : wv[dof]; };
  const T sw = AXPY ? T(a.sdev[a.sidx]) : (LOADOP == LOAD_MUL ? a.load_scale : T(0));

  // ---- loader geometry (plane independent): this thread's LSLOTS slots of the staged plane
  // image (NY rows of NZ nodes, 3 values per node, contiguous along a row), their in-plane DOF
  // offsets and validity.  A slot is a whole node, so its three adjacent values are read by one
  // wide load (single values over the flat row image, coalesced as well, made the solves 2-3 %
  // slower: three times the vector memory instructions).
  constexpr int NNODE = TL::NY * NZ;
  constexpr int LSLOTS = (NNODE + BLOCK - 1) / BLOCK;
  int ld_off[LSLOTS];
  unsigned ld_okbits = 0;
#pragma unroll
  for (int j = 0; j < LSLOTS; ++j) {
    const int k = tid + j * BLOCK;
    const int row = k / NZ;
    const int iy = ybase + row;
    const int iz = zbase + (k - row * NZ);
    const bool ok = k < (tye + 2) * NZ && iy >= 0 && iy <= ny && iz >= 0 && iz <= nz;
    ld_okbits |= unsigned(ok) << j;
    ld_off[j] = ok ? 3 * (iy * nz1 + iz) : 0;
  }
  const int e_off = col_ok ? ey * nz + ez : 0;
  const size_t pstride = size_t(3) * ny1 * nz1;     // DOFs per node plane
  const size_t estride = size_t(ny) * nz;           // elements per element plane

  // ---- branch-free prefetch of node plane ix and of this column's modulus in element plane
  // ix-1: out-of-range slots load a clamped in-bounds address, and the zeroing of those slots
  // is deferred to commit(), so no instruction consumes a loaded value before the next step
  // (the batch stays in flight through this step's compute).
  // (an FP32 operand w stays FP32 in its prefetch registers until the commit converts it)
  using WT = std::conditional_t<LOADOP == LOAD_AXPY32, float, T>;
  struct Staged {
    T v[LSLOTS][3];
    WT w[COMBINE ? LSLOTS : 1][3];
    uint8_t m[LSLOTS][3];
    unsigned keep;
    T E;
    bool E_ok;
  };
  auto prefetch = [&](Staged& st, int ix) {
    // (the x = 0 clamp is applied to whole faces in face(), so the staged plane keeps the
    //  unmasked input the epilogue needs at fixed DOFs)
    const bool plane_ok = ix >= 0 && ix <= nx;
    const size_t base = (ix >= 0 && ix <= nx) ? size_t(ix) * pstride : 0;
    unsigned keep = 0;
#pragma unroll
    for (int j = 0; j < LSLOTS; ++j) {
      // (the last slot row is partly beyond the image: waves wholly past it skip the load)
      if (j < LSLOTS - 1 || tid + j * BLOCK < NNODE) {
        const size_t dof = base + ld_off[j];
#pragma unroll
        for (int c = 0; c < 3; ++c) st.v[j][c] = load_as<S, T>(u, dof + c);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          if (COMBINE) st.w[j][c] = LOADOP == LOAD_AXPY32 ? WT(wf[dof + c]) : WT(wv[dof + c]);
          if (MASK == MASK_ARRAY) st.m[j][c] = a.mask[dof + c];
        }
      }
      keep |= unsigned(plane_ok && ((ld_okbits >> j) & 1u)) << j;
    }
    st.keep = keep;
    const int ex = ix - 1;
    st.E_ok = col_ok && ex >= 0 && ex < nx;
    st.E = E[(st.E_ok ? size_t(ex) * estride : 0) + e_off];
  };
  auto commit = [&](const Staged& st, int ix) {
    // the prefetched modulus goes through LDS like the plane: a store before the barrier cannot
    // be sunk below the next prefetch, so its register is free to receive the next load (kept in
    // a register, the compiler may keep it live past the next load and then copy it, waiting for
    // every load in flight)
    sE[tid] = st.E_ok ? st.E : T(0);
    T (*pl)[PLANE] = su[ix & 1];
#pragma unroll
    for (int j = 0; j < LSLOTS; ++j) {
      const int k = tid + j * BLOCK;
      if (j < LSLOTS - 1 || k < NNODE) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          bool keep = (st.keep >> j) & 1u;
          if (MASK == MASK_ARRAY) keep = keep && st.m[j][c] != 0;
          const T v = AXPY ? st.v[j][c] + sw * T(st.w[j][c])
                      : (LOADOP == LOAD_MUL ? sw * (st.v[j][c] * T(st.w[j][c])) : st.v[j][c]);
          pl[c][k] = keep ? v : T(0);
        }
      }
    }
  };
  // 2x2 (y,z) transform of this column's face of node plane `plane`: F[q], q = py + 2 pz.
  auto face = [&](T (&F)[4][3], int plane) {
    const int i00 = ly * NZ + lz;
    const int pb = plane & 1;
    T w[3][4];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      w[c][0] = su[pb][c][i00];
      w[c][1] = su[pb][c][i00 + 1];
      w[c][2] = su[pb][c][i00 + NZ];
      w[c][3] = su[pb][c][i00 + NZ + 1];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T s0 = w[c][0] + w[c][1], d0 = w[c][0] - w[c][1];
      const T s1 = w[c][2] + w[c][3], d1 = w[c][2] - w[c][3];
      F[0][c] = s0 + s1;
      F[1][c] = s0 - s1;
      F[2][c] = d0 + d1;
      F[3][c] = d0 - d1;
    }
    if (MASK == MASK_X0 && plane == 0) {     // every DOF of node plane 0 is fixed: M u = 0 there
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) F[q][c] = T(0);
    }
  };

  T Fl[4][3], Fr[4][3], Gc[4][3];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int c = 0; c < 3; ++c) Gc[q][c] = T(0);

  double dot_acc = 0.0, dot_acc2 = 0.0;
  T Eprev = T(0);                  // modulus of this column in element plane ex - 1

  // ---- epilogue operands of this thread's owned node: b, the previous iterate / d, dinv (and u
  // itself with a mask array); otherwise the input u at the node is read from the staged plane
  // (still in LDS: planes are double-buffered)
  constexpr bool NEED_B = MODE == MODE_RESID || NEEDS_DINV;
  constexpr bool NEED_U = MODE != MODE_RAW || MASK != MASK_NONE || COMBINE;
  // (issued at the start of step ex, they land during the element compute)
  T ep_b[3], ep_p[3], ep_di[3], ep_u[3], ep_z[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) ep_b[c] = ep_p[c] = ep_di[c] = ep_u[c] = ep_z[c] = T(0);
  const int jn_own = (ly + 1) * NZ + (lz + 1);
  // the element columns (ly+1, lz+1), (ly+1, lz), (ly, lz+1) around the owned node besides this
  // thread's own (ly, lz), as thread numbers; past the last column of a row (lz + 1 == EZ) there is
  // none: the zero pad
  const bool zlast = lz + 1 >= EZ;
  const int w0 = zlast ? SCN - 1 : tid + EZ + 1, w1 = tid + EZ, w2 = zlast ? SCN - 1 : tid + 1;
  // (each operand's three components are read by adjacent loads of one basic block, so they merge
  //  into one wide load per node; the previous Chebyshev iterate is read unconditionally and
  //  discarded when the step does not use it, except by the first step from a zero iterate (the
  //  MUL load), which never uses it)
  const T* __restrict__ prev_src = (MODE == MODE_CHEB3 || MODE == MODE_CHEB3D) ? (a.prev ? a.prev : a.y) : nullptr;
  const bool prev_used = a.cheb_c != T(0) && !(a.flags & MV_PREV_ZERO);
  auto fetch_epilogue = [&](int ix) {
    const bool use = node_own && ix >= ox0 && ix < ox1;
    const size_t node = use ? (size_t(ix) * ny1 + iy_own) * nz1 + iz_own : 0;
    if (NEED_U && MASK == MASK_ARRAY) {         // the staged plane is masked: fetch the input
#pragma unroll
      for (int c = 0; c < 3; ++c) ep_u[c] = load_as<S, T>(u, 3 * node + c);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (AXPY) ep_u[c] += sw * w_at(3 * node + c);
        if (LOADOP == LOAD_MUL) ep_u[c] = sw * (ep_u[c] * wv[3 * node + c]);
      }
    }
    if (NEED_B) {
#pragma unroll
      for (int c = 0; c < 3; ++c) ep_b[c] = a.b[3 * node + c];
    }
    if (MODE == MODE_CHEB) {
#pragma unroll
      for (int c = 0; c < 3; ++c) ep_p[c] = a.d[3 * node + c];
    }
    if ((MODE == MODE_CHEB3 || MODE == MODE_CHEB3D) && LOADOP != LOAD_MUL) {
#pragma unroll
      for (int c = 0; c < 3; ++c) ep_p[c] = prev_src[3 * node + c];
#pragma unroll
      for (int c = 0; c < 3; ++c) ep_p[c] = prev_used ? ep_p[c] : T(0);
    }
    if (MODE == MODE_CHEB3D) {
#pragma unroll
      for (int c = 0; c < 3; ++c) ep_z[c] = a.zold[3 * node + c];
    }
    if (NEEDS_DINV && !dinv_from_E) {
#pragma unroll
      for (int c = 0; c < 3; ++c) ep_di[c] = a.dinv[3 * node + c];
    }
  };

  // Prologue: planes ox0-1 and ox0 (and the modulus of element ox0 - 1) loaded together, so the
  // launch waits for one memory latency before its first step, not two.  The first commit of the
  // loop writes the other plane buffer, and the corner slots are first written two barriers on.
  Staged st;
  {
    Staged st0;
    prefetch(st0, ox0 - 1);
    prefetch(st, ox0);
    commit(st0, ox0 - 1);
  }
  __syncthreads();                 // plane ox0 - 1 staged
  face(Fl, ox0 - 1);

  // Per step: commit plane ex+1 | barrier | prefetch ex+2 | face + element ex | contributions
  // | barrier | gather.  The second barrier also orders this step's face() reads before the
  // next step's commit, and the gather before the next step's contribution writes.
  for (int ex = ox0 - 1; ex < ox1; ++ex) {
    commit(st, ex + 1);
    __syncthreads();
    const T Ee = sE[tid];
    // this step's epilogue operands first: loads complete in issue order, so the gather's wait
    // for them then leaves the prefetch of plane ex + 2 in flight
    fetch_epilogue(ex);
    __builtin_amdgcn_sched_barrier(0);
    prefetch(st, ex + 2);
    if (RECOMPUTE_FL) face(Fl, ex);
    face(Fr, ex + 1);

    T Gl[4][3], Gr[4][3];
    {
      T v[8][3];
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          v[2 * q][c] = Fl[q][c] + Fr[q][c];
          v[2 * q + 1][c] = Fl[q][c] - Fr[q][c];
        }
      T f[8][3];
      element_blocks<T, CUBE>(v, Ee, a, f);   // Ee == 0 outside the grid: f == 0
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          Gl[q][c] = f[2 * q][c] + f[2 * q + 1][c];
          Gr[q][c] = f[2 * q][c] - f[2 * q + 1][c];
        }
    }

    const bool finalize = ex >= ox0;     // node plane ex complete: carried right half + left half
    T own[3];                            // corner (1, 1): this thread's owned node (ly+1, lz+1)
    if (finalize && col_in_tile) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T g0 = Gc[0][c] + Gl[0][c], g1 = Gc[1][c] + Gl[1][c];
        const T g2 = Gc[2][c] + Gl[2][c], g3 = Gc[3][c] + Gl[3][c];
        const T A0 = g0 + g1, A1 = g0 - g1, B0 = g2 + g3, B1 = g2 - g3;
        sc[0][c][tid] = A0 + B0;                       // corner (b=0, c=0): node (ly, lz)
        sc[1][c][tid] = A0 - B0;                       // corner (0, 1): node (ly, lz+1)
        sc[2][c][tid] = A1 + B1;                       // corner (1, 0): node (ly+1, lz)
        own[c] = A1 - B1;                              // corner (1, 1): node (ly+1, lz+1)
      }
      if (dinv_from_E) se[tid] = Eprev + Ee;
    }
    // unmasked input at the owned node of plane ex, read before the barrier that lets the next
    // step overwrite this plane buffer
    if (NEED_U && MASK != MASK_ARRAY && finalize && node_own) {
#pragma unroll
      for (int c = 0; c < 3; ++c) ep_u[c] = su[ex & 1][c][jn_own];
    }
    __syncthreads();
    if (finalize && node_own) {
      const size_t node = (size_t(ex) * ny1 + iy_own) * nz1 + iz_own;
      T part[3][4];      // corner k of the node comes from column w_k
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        part[c][0] = sc[0][c][w0];
        part[c][1] = sc[1][c][w1];
        part[c][2] = sc[2][c][w2];
        part[c][3] = own[c];
      }
      T rsum = T(0);     // 1 / (sum of the moduli of the node's incident elements)
      if (dinv_from_E) { // incident columns (ly|ly+1, lz|lz+1) of the owned node (ly+1, lz+1)
        const T esum = ((Eprev + Ee + se[w1]) + se[w2]) + se[w0];
        rsum = reciprocal(esum);
      }
      T outv[3], dv[3] = {T(0), T(0), T(0)};
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T ku = ((part[c][0] + part[c][1]) + part[c][2]) + part[c][3];
        const size_t dof = 3 * node + c;
        bool free_dof = true;
        if (MASK == MASK_X0) free_dof = ex > 0;
        if (MASK == MASK_ARRAY) free_dof = a.mask[dof] != 0;
        const T uo = ep_u[c];           // unmasked input at the owned DOF
        // constrained operator (M K M + I - M) u, or the raw K u without a mask
        const T au = (MASK == MASK_NONE || free_dof) ? ku : (a.fixed_zero ? T(0) : uo);
        T out;
        if (MODE == MODE_RAW) {
          out = au;
        } else if (MODE == MODE_RESID) {
          out = ep_b[c] - au;
          if (a.partial) dot_acc += double(out) * double(out);
        } else if (NEEDS_DINV) {
          T dinv;
          if (dinv_from_E) {
            const T kdi = c == 0 ? a.kdi[0] : (c == 1 ? a.kdi[1] : a.kdi[2]);
            dinv = free_dof ? kdi * rsum : T(1);
          }
          else dinv = ep_di[c];
          const T z = dinv * (ep_b[c] - au);
          if (MODE == MODE_JACOBI) {
            out = uo + a.omega * z;
          } else if (MODE == MODE_CHEB) {
            const T d = a.cheb_a * (a.cheb_c == T(0) ? z : z + a.cheb_c * ep_p[c]);
            dv[c] = d;
            out = uo + d;
          } else {
            out = uo + a.cheb_a * (a.cheb_c == T(0) ? z : z + a.cheb_c * (uo - ep_p[c]));
          }
        } else {  // MODE_DOT: y = A u and sum u . A u
          out = au;
          dot_acc += double(uo) * double(au);
        }
        if (MASK != MASK_NONE && a.fixed_zero && !free_dof) out = T(0);
        if (MODE == MODE_CHEB3D) {     // r . z_new and r . (z_new - z_old) of the Krylov update
          const double rb = double(ep_b[c]), zo = double(out);
          dot_acc += rb * zo;
          dot_acc2 += rb * (zo - double(ep_z[c]));
        }
        outv[c] = out;
      }
      // each output's three components by adjacent stores (one wide store per node)
      const size_t d0 = 3 * node;
      if (MODE == MODE_CHEB) {
#pragma unroll
        for (int c = 0; c < 3; ++c) a.d[d0 + c] = dv[c];
      }
      if ((MODE != MODE_DOT && MODE != MODE_RESID) || a.y) {   // DOT / RESID: y optional
#pragma unroll
        for (int c = 0; c < 3; ++c) a.y[d0 + c] = outv[c];
      }
      if (COMBINE) {
#pragma unroll
        for (int c = 0; c < 3; ++c) a.u_out[d0 + c] = ep_u[c];
      }
      if (MODE == MODE_RESID && a.y32) {
#pragma unroll
        for (int c = 0; c < 3; ++c) a.y32[d0 + c] = float(outv[c]);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Gc[q][c] = Gr[q][c];
        if (!RECOMPUTE_FL) Fl[q][c] = Fr[q][c];
      }
    Eprev = Ee;
  }

  if (MODE == MODE_DOT || (MODE == MODE_RESID && a.partial != nullptr)) {
    const double s = block_sum<double, BLOCK>(dot_acc, red);
    if (tid == 0) a.partial[logical] = s;
  }
  if constexpr (MODE == MODE_CHEB3D) {
    __shared__ double red2[BLOCK / 64];
    const double s0 = block_sum<double, BLOCK>(dot_acc, red);
    const double s1 = block_sum<double, BLOCK>(dot_acc2, red2);
    if (tid == 0) {
      a.partial[logical] = s0;
      a.partial[a.n_tiles + logical] = s1;
    }
  }
}

// Resident waves per SIMD the register allocation must allow: 3 for FP64 (at most 168 VGPRs), 4 for FP32
// (at most 128).  A workgroup of 5 waves (the 8 x 32 and 16 x 16 tiles) would otherwise let the compiler
// use up to 256 VGPRs, and past 168 only one such workgroup fits on a CU instead of two.
template <typename T> constexpr int kMinWaves = sizeof(T) == 8 ? 3 : 4;

template <typename T, typename S, int MODE, int MASK, typename TL, bool CUBE, int LOADOP = LOAD_NONE>
__global__ void __launch_bounds__(TL::BLOCK, kMinWaves<T>) parity_xmarch(MatvecParams<T> a) {
  const int logical = xcd_logical(blockIdx.x, gridDim.x);
  if (logical >= a.n_tiles) return;
  xmarch_tile<T, S, MODE, MASK, TL, CUBE, LOADOP>(a, logical);
}

// Batched product: problem i owns the logical tiles [offs[i], offs[i+1]) of one launch and
// reads its parameters (pointers, grid, plan, epilogue scalars, coefficients) from descs[i].
template <typename T, typename S, int MODE, int MASK, typename TL, bool CUBE, int LOADOP = LOAD_NONE>
__global__ void __launch_bounds__(TL::BLOCK, kMinWaves<T>) parity_xmarch_batch(const MatvecParams<T>* __restrict__ descs,
                                                                 const int* __restrict__ offs, int n) {
  const int logical = xcd_logical(blockIdx.x, gridDim.x);
  if (logical >= offs[n]) return;
  int lo = 0, hi = n;                    // offs[lo] <= logical < offs[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (offs[mid] <= logical) lo = mid; else hi = mid;
  }
  lo = __builtin_amdgcn_readfirstlane(lo);   // uniform: the descriptor is read with scalar loads
  xmarch_tile<T, S, MODE, MASK, TL, CUBE, LOADOP>(descs[lo], logical - offs[lo]);
}

// ------------------------------------------------------------------------------ launcher
using Tile16 = Tile<16, 32>;
using Tile8 = Tile<8, 32>;
using Tile4 = Tile<4, 8>;
using Tile16s = Tile<16, 16>;     // batched products only
// one workgroup per CU, its waves spread evenly over the four SIMDs: FP64 (<= 168 VGPRs) takes
// three waves per SIMD (23 x 33 = 759 element columns), FP32 (<= 128 VGPRs) four (31 x 33 = 1023)
template <typename T>
using TileWide = std::conditional_t<sizeof(T) == 8, Tile<22, 32>, Tile<30, 32>>;

// (the cached values are initialised once, thread-safely, on the first call)
static int device_cus() {
  static const int cus = [] {
    int dev = 0, n = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      n = prop.multiProcessorCount;
    return n > 0 ? n : 256;
  }();
  return cus;
}

// Resident workgroups per CU of one kernel instantiation (cached).
template <typename T, typename S, typename TL, bool CUBE>
static int resident_per_cu() {
  static const int occ = [] {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, parity_xmarch<T, S, MODE_RAW, MASK_NONE, TL, CUBE>,
                                                    TL::BLOCK, 0) != hipSuccess || n <= 0)
      n = 1;
    return n;
  }();
  return occ;
}

// x chunking: with P tiles per plane, S = CUs * resident workgroups per CU and c chunks, the
// sweep takes about ceil(P c / S) rounds of (nx / c + 1) plane steps (the +1 is the halo
// element plane recomputed at each chunk start); pick the c that minimises it.
static void plan_chunks(int nx, int per_plane, int slots, int lx_hint, int* lx, int* chunks) {
  int l = lx_hint;
  if (l <= 0) {
    double best = 1e300;
    l = nx;
    for (int c = 1; c <= nx; ++c) {
      const int lc = (nx + c - 1) / c;
      const int cc = (nx + lc - 1) / lc;
      const double rounds = double((long(per_plane) * cc + slots - 1) / slots);
      const double cost = rounds * (lc + 1);
      if (cost < best - 1e-9) { best = cost; l = lc; }
    }
  }
  if (l > nx) l = nx;
  if (l < 1) l = 1;
  *lx = l;
  *chunks = (nx + l - 1) / l;
}

template <typename TL>
static void plan_tile(int nx, int ny, int nz, int lx_hint, int occ, MatvecPlan* pl) {
  const int ty = (ny + TL::TY - 1) / TL::TY;
  const int tz = (nz + TL::TZ - 1) / TL::TZ;
  // the ny element rows spread evenly over the y tiles: each owns tile_ty <= TY node rows (the
  // last one tile_ty + 1 at most, the extra node row at y = ny)
  pl->tile_ty = (ny + ty - 1) / ty;
  int lx = 0, cx = 0;
  plan_chunks(nx, ty * tz, device_cus() * occ, lx_hint, &lx, &cx);
  pl->tiles_y = ty;
  pl->tiles_z = tz;
  pl->chunks_x = cx;
  pl->lx = lx;
  pl->n_tiles = ty * tz * cx;
  pl->grid = ((pl->n_tiles + 7) / 8) * 8;
  // relative cost: rounds x plane steps x waves on the busiest SIMD of a CU.  The workgroups of a
  // CU step in lockstep through their barriers, so a step lasts as long as the SIMD holding the
  // most waves takes to issue them (9 waves: 3 + 2 + 2 + 2, the SIMD with 3 sets the pace).
  const long slots = long(device_cus()) * occ;
  const long rounds = (pl->n_tiles + slots - 1) / slots;
  const int waves = occ * (TL::BLOCK / 64);
  pl->cost = double(rounds) * (lx + 1) * ((waves + 3) / 4);
}

template <typename T, typename S, bool CUBE>
static MatvecPlan plan_typed(int nx, int ny, int nz, int lx_hint, int tile_hint) {
  MatvecPlan a{}, b{};
  plan_tile<Tile16>(nx, ny, nz, lx_hint, resident_per_cu<T, S, Tile16, CUBE>(), &a);
  a.tile = MV_TILE_16x32;
  plan_tile<Tile8>(nx, ny, nz, lx_hint, resident_per_cu<T, S, Tile8, CUBE>(), &b);
  b.tile = MV_TILE_8x32;
  if (tile_hint == MV_TILE_16x32) return a;
  if (tile_hint == MV_TILE_8x32) return b;
  if (tile_hint == MV_TILE_4x8) {
    MatvecPlan c{};
    plan_tile<Tile4>(nx, ny, nz, lx_hint, resident_per_cu<T, S, Tile4, CUBE>(), &c);
    c.tile = MV_TILE_4x8;
    return c;
  }
  MatvecPlan c{};
  plan_tile<TileWide<T>>(nx, ny, nz, lx_hint, resident_per_cu<T, S, TileWide<T>, CUBE>(), &c);
  c.tile = MV_TILE_WIDE;
  if (tile_hint == MV_TILE_WIDE) return c;
  MatvecPlan best = (b.cost < a.cost) ? b : a;
  return c.cost < best.cost ? c : best;
}

#if defined(VH_MATVEC_F64)
MatvecPlan matvec_plan_f64(bool cube, int nx, int ny, int nz, int lx_hint, int tile_hint) {
  return cube ? plan_typed<double, double, true>(nx, ny, nz, lx_hint, tile_hint)
              : plan_typed<double, double, false>(nx, ny, nz, lx_hint, tile_hint);
}
#elif defined(VH_MATVEC_F32)
MatvecPlan matvec_plan_f32(int storage, bool cube, int nx, int ny, int nz, int lx_hint, int tile_hint) {
  if (storage == BF16)
    return cube ? plan_typed<float, uint16_t, true>(nx, ny, nz, lx_hint, tile_hint)
                : plan_typed<float, uint16_t, false>(nx, ny, nz, lx_hint, tile_hint);
  return cube ? plan_typed<float, float, true>(nx, ny, nz, lx_hint, tile_hint)
              : plan_typed<float, float, false>(nx, ny, nz, lx_hint, tile_hint);
}
#endif

template <typename T, typename S, int MODE, int MASK, bool CUBE, int LOP>
static void launch_tiles(const MatvecParams<T>& p, int grid, hipStream_t st) {
  if (p.tile == MV_TILE_8x32)
    hipLaunchKernelGGL((parity_xmarch<T, S, MODE, MASK, Tile8, CUBE, LOP>), dim3(grid), dim3(Tile8::BLOCK), 0, st, p);
  else if (p.tile == MV_TILE_4x8)
    hipLaunchKernelGGL((parity_xmarch<T, S, MODE, MASK, Tile4, CUBE, LOP>), dim3(grid), dim3(Tile4::BLOCK), 0, st, p);
  else if (p.tile == MV_TILE_WIDE)
    hipLaunchKernelGGL((parity_xmarch<T, S, MODE, MASK, TileWide<T>, CUBE, LOP>), dim3(grid), dim3(TileWide<T>::BLOCK), 0,
                       st, p);
  else
    hipLaunchKernelGGL((parity_xmarch<T, S, MODE, MASK, Tile16, CUBE, LOP>), dim3(grid), dim3(Tile16::BLOCK), 0, st, p);
}

template <typename T, typename S, int MODE, int MASK>
static void launch_t(const MatvecParams<T>& p, int grid, hipStream_t st) {
  // load-time combination (voxel-cube blocks only; checked by the caller): AXPY for the Krylov
  // passes, MUL for the first Chebyshev step
  constexpr int LOP = (MODE == MODE_DOT || MODE == MODE_RESID) ? LOAD_AXPY : (MODE == MODE_CHEB3 ? LOAD_MUL : LOAD_NONE);
  if constexpr (LOP != LOAD_NONE) {
    if (p.w) {
      // an FP32 w in the FP64 RESID pass (checked by the caller)
      if constexpr (sizeof(T) == 8 && MODE == MODE_RESID) {
        if (p.flags & MV_W_F32) {
          launch_tiles<T, S, MODE, MASK, true, LOAD_AXPY32>(p, grid, st);
          return;
        }
      }
      launch_tiles<T, S, MODE, MASK, true, LOP>(p, grid, st);
      return;
    }
  }
  if (p.cube) launch_tiles<T, S, MODE, MASK, true, LOAD_NONE>(p, grid, st);
  else launch_tiles<T, S, MODE, MASK, false, LOAD_NONE>(p, grid, st);
}

template <int MASK>
static void launch_dot_f32_input(const MatvecParams<double>& p, int grid, hipStream_t st) {
  launch_tiles<double, float, MODE_DOT, MASK, true, LOAD_AXPY>(p, grid, st);
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
    case MODE_CHEB3: launch_mask<T, S, MODE_CHEB3>(p, mask_kind, grid, st); break;
    case MODE_CHEB3D: launch_mask<T, S, MODE_CHEB3D>(p, mask_kind, grid, st); break;
    default: launch_mask<T, S, MODE_DOT>(p, mask_kind, grid, st); break;
  }
}

template <typename T>
int matvec_launch(MatvecParams<T> p, int storage, int mode, int mask_kind, hipStream_t st) {
  const int dtype = sizeof(T) == 8 ? F64 : F32;
  const MatvecPlan pl = matvec_plan(dtype, storage, p.cube != 0, p.nx, p.ny, p.nz, p.lx, p.tile);
  p.tile = pl.tile;
  p.lx = pl.lx;
  p.tiles_y = pl.tiles_y;
  p.tiles_z = pl.tiles_z;
  p.tile_ty = pl.tile_ty;
  p.chunks_x = pl.chunks_x;
  p.n_tiles = pl.n_tiles;
  if (storage == BF16) {
    if constexpr (sizeof(T) == 4) launch_mode<T, uint16_t>(p, mode, mask_kind, pl.grid, st);
    else return fail(-2, "bf16 storage requires fp32 compute");
  } else if (storage == F32 && sizeof(T) == 8) {
    // FP32 input of an FP64 pass: the DOT pass of the mixed-precision Krylov loop, which forms
    // p = z + beta p_old from the FP32 preconditioned residual z (voxel-cube blocks, AXPY load)
    if constexpr (sizeof(T) == 8) {
      if (mode != MODE_DOT || !p.w || !p.cube) return fail(-2, "fp32 input of an fp64 pass: DOT with AXPY load only");
      switch (mask_kind) {
        case MASK_NONE: launch_dot_f32_input<MASK_NONE>(p, pl.grid, st); break;
        case MASK_ARRAY: launch_dot_f32_input<MASK_ARRAY>(p, pl.grid, st); break;
        default: launch_dot_f32_input<MASK_X0>(p, pl.grid, st); break;
      }
    }
  } else {
    launch_mode<T, T>(p, mode, mask_kind, pl.grid, st);
  }
  return hip_status(hipGetLastError(), "parity_xmarch launch");
}

#if defined(VH_MATVEC_F64)
template int matvec_launch<double>(MatvecParams<double>, int, int, int, hipStream_t);
#elif defined(VH_MATVEC_F32)
template int matvec_launch<float>(MatvecParams<float>, int, int, int, hipStream_t);
#endif

// ------------------------------------------------------------------------ batched launcher
// Instantiated for the voxel-cube coefficients and the passes of the batched solvers: RAW,
// RESID and DOT (with or without the AXPY load), JACOBI, CHEB3 (with or without the MUL load),
// masks X0 and ARRAY, both tiles.
#if defined(VH_MATVEC_BATCH_F64) || defined(VH_MATVEC_BATCH_F32)
template <typename T, int MODE, int MASK, int LOADOP>
static void batch_tile(const MatvecBatch& bt, hipStream_t st) {
  const auto* d = static_cast<const MatvecParams<T>*>(bt.descs);
  const int grid = ((bt.total + 7) / 8) * 8;
  if (bt.tile == MV_TILE_8x32)
    hipLaunchKernelGGL((parity_xmarch_batch<T, T, MODE, MASK, Tile8, true, LOADOP>), dim3(grid), dim3(Tile8::BLOCK), 0, st,
                       d, bt.offs, bt.n);
  else if (bt.tile == MV_TILE_16x16)
    hipLaunchKernelGGL((parity_xmarch_batch<T, T, MODE, MASK, Tile16s, true, LOADOP>), dim3(grid), dim3(Tile16s::BLOCK),
                       0, st, d, bt.offs, bt.n);
  else
    hipLaunchKernelGGL((parity_xmarch_batch<T, T, MODE, MASK, Tile16, true, LOADOP>), dim3(grid), dim3(Tile16::BLOCK), 0,
                       st, d, bt.offs, bt.n);
}

template <typename T, int MASK>
static int batch_mode(const MatvecBatch& bt, hipStream_t st) {
  switch (bt.mode) {
    case MODE_RAW: batch_tile<T, MODE_RAW, MASK, LOAD_NONE>(bt, st); break;
    case MODE_RESID:
      if constexpr (sizeof(T) == 8) {
        if (bt.combine == 2) {          // FP32 w (checked at creation)
          batch_tile<T, MODE_RESID, MASK, LOAD_AXPY32>(bt, st);
          break;
        }
      }
      if (bt.combine) batch_tile<T, MODE_RESID, MASK, LOAD_AXPY>(bt, st);
      else batch_tile<T, MODE_RESID, MASK, LOAD_NONE>(bt, st);
      break;
    case MODE_DOT:
      if (bt.combine) batch_tile<T, MODE_DOT, MASK, LOAD_AXPY>(bt, st);
      else batch_tile<T, MODE_DOT, MASK, LOAD_NONE>(bt, st);
      break;
    case MODE_JACOBI: batch_tile<T, MODE_JACOBI, MASK, LOAD_NONE>(bt, st); break;
    case MODE_CHEB3:
      if (bt.combine) batch_tile<T, MODE_CHEB3, MASK, LOAD_MUL>(bt, st);
      else batch_tile<T, MODE_CHEB3, MASK, LOAD_NONE>(bt, st);
      break;
    default: return fail(-2, "batched product: mode %d not available", bt.mode);
  }
  return hip_status(hipGetLastError(), "parity_xmarch_batch launch");
}

template <typename T>
int matvec_batch_launch(const MatvecBatch& bt, hipStream_t st) {
  if (bt.mask_kind == MASK_X0) return batch_mode<T, MASK_X0>(bt, st);
  if (bt.mask_kind == MASK_ARRAY) return batch_mode<T, MASK_ARRAY>(bt, st);
  return fail(-2, "batched product: mask kind %d not available", bt.mask_kind);
}

// occupancy of the batched kernel itself (the single-problem kernels live in other translation
// units), taken for the Chebyshev smoothing pass that makes up most of a batched solve (it holds
// more registers than the raw product)
template <typename T, typename TL>
static int resident_batch() {
  static const int occ = [] {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, parity_xmarch_batch<T, T, MODE_CHEB3, MASK_X0, TL, true, LOAD_NONE>,
                                                    TL::BLOCK, 0) != hipSuccess || n <= 0)
      n = 1;
    return n;
  }();
  return occ;
}

template <typename T>
MatvecPlan matvec_batch_plan(int nx, int ny, int nz, int lx_hint, int tile) {
  MatvecPlan p{};
  if (tile == MV_TILE_8x32) {
    plan_tile<Tile8>(nx, ny, nz, lx_hint, resident_batch<T, Tile8>(), &p);
    p.tile = MV_TILE_8x32;
  } else if (tile == MV_TILE_16x16) {
    plan_tile<Tile16s>(nx, ny, nz, lx_hint, resident_batch<T, Tile16s>(), &p);
    p.tile = MV_TILE_16x16;
  } else {
    plan_tile<Tile16>(nx, ny, nz, lx_hint, resident_batch<T, Tile16>(), &p);
    p.tile = MV_TILE_16x32;
  }
  return p;
}

// The problems of one batched launch share the chip: planned one by one, each small problem would
// be cut into one-element x chunks to fill every CU by itself, recomputing a halo element plane
// per owned one.  Here one chunk length L is chosen for the whole batch, minimising the list-
// scheduling bound  (sum of workgroup plane steps) / (resident workgroup slots) + (longest
// workgroup): a workgroup of a chunk of l planes takes l + 2 steps (the halo element plane and
// the first node plane loaded before the march).  Problem i gets ceil(nx_i / L) chunks, spread
// evenly over its nx_i element planes.
template <typename T>
void matvec_batch_chunks(int n, const int* nx, const int* ny, const int* nz, int tile, int* lx) {
  if (n <= 0) return;
  std::vector<long> per(n);
  int max_nx = 1;
  for (int i = 0; i < n; ++i) {
    const MatvecPlan p = matvec_batch_plan<T>(nx[i], ny[i], nz[i], nx[i], tile);
    per[i] = long(p.tiles_y) * p.tiles_z;
    max_nx = std::max(max_nx, nx[i]);
  }
  const int occ = tile == MV_TILE_8x32    ? resident_batch<T, Tile8>()
                  : tile == MV_TILE_16x16 ? resident_batch<T, Tile16s>()
                                          : resident_batch<T, Tile16>();
  const double slots = double(device_cus()) * occ;
  int best_l = max_nx;
  double best = 1e300;
  for (int L = 1; L <= max_nx; ++L) {
    double work = 0.0;
    int longest = 0;
    for (int i = 0; i < n; ++i) {
      const int c = (nx[i] + L - 1) / L, l = (nx[i] + c - 1) / c;
      work += double(per[i]) * c * (l + 2);
      longest = std::max(longest, l + 2);
    }
    const double cost = work / slots + longest;
    if (cost < best - 1e-9) {
      best = cost;
      best_l = L;
    }
  }
  for (int i = 0; i < n; ++i) {
    const int c = (nx[i] + best_l - 1) / best_l;
    lx[i] = (nx[i] + c - 1) / c;
  }
}
#endif
#if defined(VH_MATVEC_BATCH_F64)
template int matvec_batch_launch<double>(const MatvecBatch&, hipStream_t);
template MatvecPlan matvec_batch_plan<double>(int, int, int, int, int);
template void matvec_batch_chunks<double>(int, const int*, const int*, const int*, int, int*);
#elif defined(VH_MATVEC_BATCH_F32)
template int matvec_batch_launch<float>(const MatvecBatch&, hipStream_t);
template MatvecPlan matvec_batch_plan<float>(int, int, int, int, int);
template void matvec_batch_chunks<float>(int, const int*, const int*, const int*, int, int*);
#endif

}  // namespace vh
