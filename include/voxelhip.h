/* voxelhip: MI355X (gfx950) matrix-free Hex8 voxel elasticity — C ABI.
 *
 * Every entry point returns 0 on success and a negative code on failure (-1 HIP error,
 * -2 bad argument, -3 outside the parity-block model); vh_last_error() gives the message of
 * the calling thread's last failure.  Array arguments are DEVICE pointers (plain, contiguous)
 * unless stated; `stream` is a hipStream_t (NULL = the null stream).  Launches are
 * asynchronous on `stream`; nothing synchronises the device except vh_modal_blocks (host only).
 *
 * Layout ("author" layout of the reference, src/voxel_mg.py:20-21,32-38): element
 * e = (ex*ny + ey)*nz + ez, node n = (ix*(ny+1) + iy)*(nz+1) + iz, DOF 3n + c.
 * dtype: 0 = float64, 1 = float32; storage 2 = bfloat16 input vector with float32 compute.
 * mask_kind: 0 = none (raw product), 1 = per-DOF uint8 array (1 = free), 2 = the x = 0 clamp
 * (every DOF of node plane ix = 0 fixed; no array needed).
 */
#ifndef VOXELHIP_H
#define VOXELHIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VH_ABI_VERSION 4

enum { VH_F64 = 0, VH_F32 = 1, VH_BF16 = 2 };
enum { VH_MASK_NONE = 0, VH_MASK_ARRAY = 1, VH_MASK_X0 = 2 };
enum { VH_MODE_RAW = 0, VH_MODE_RESID = 1, VH_MODE_JACOBI = 2, VH_MODE_CHEB = 3, VH_MODE_DOT = 4,
       VH_MODE_CHEB3 = 5, VH_MODE_CHEB3D = 6 };
enum { VH_RED_DOT = 0, VH_RED_DOT2 = 1, VH_RED_DOTDIFF = 2, VH_RED_NORM2 = 3 };
enum { VH_OP_NONE = 0, VH_OP_ALPHA = 1, VH_OP_BETA_FR = 2, VH_OP_BETA_PR = 3, VH_OP_SET_RZ = 4,
       VH_OP_INVSQRT = 5 /* scal[dst] = 1 / sqrt(sum): a device-side normalisation factor */,
       VH_OP_RESNORM = 6 /* scal[dst] = |r|^2 of an iteration; unless VH_S_DONE: VH_S_ITS += 1 and
                            VH_S_DONE = 1 once |r|^2 <= scal[VH_S_TOL2] (convergence on the device,
                            so the host can queue iterations ahead without reading each one) */ };
/* device scalar slots used by vh_finish / vh_cg_update / vh_xpby; VH_S_DONE != 0 freezes the
 * recurrence (alpha = beta = 0), for converged problems of a batch */
enum { VH_S_RZ = 0, VH_S_PQ = 1, VH_S_ALPHA = 2, VH_S_BETA = 3, VH_S_FLAG = 4, VH_S_DONE = 10, VH_S_TOL2 = 11,
       VH_S_ITS = 12 };

int vh_abi_version(void);
const char* vh_last_error(void);

/* Parity-block factorisation of a 24x24 element matrix (host pointers): blocks72 = the 8
 * 3x3 blocks of C = T ke T^T / 64, block g coupling (mode g^1, x), (mode g^2, y), (mode g^4, z).
 * Returns -3 if the Frobenius share outside the blocks exceeds `tolerance` (not an
 * axis-aligned isotropic brick).
 * Replaces: Modal(ke) code/hex_modal/modal.py:55-69 (host setup of the kernel's C[72]). */
int vh_modal_blocks(const double* ke576, double tolerance, double* blocks72, double* discarded);

/* y = K(E) u (mask_kind 0), or the constrained A u = M K M u + (I - M) u.
 * blocks72 is a HOST pointer (FP64); E, u, y device arrays of `dtype` (u bf16 if
 * storage == VH_BF16).  y is fully overwritten; no zeroing needed; y must not alias u.
 * Replaces: modal8 / modal8_muladd kernels code/hex_modal_cuda/cuda.py:18-45,
 * experiments/session_e/dependencies/cuda.py:87-108 with CudaModal.matvec_full
 * code/hex_modal_cuda/cuda.py:96-101; NumpyProduct.raw src/voxel_mg.py:50-53;
 * Level.apply src/voxel_mg.py:129-130. */
int vh_matvec(int dtype, int storage, const double* blocks72, const void* E, const void* u, void* y,
              int nx, int ny, int nz, int mask_kind, const uint8_t* mask, void* stream);

/* Fused epilogues over the same single pass (A as in vh_matvec; with a mask, fixed_zero makes
 * the output 0 at fixed DOFs, i.e. M K M u for RAW/DOT and (b - A u) M for RESID):  RESID y = b - A u;  JACOBI y = u + omega dinv (b - A u);
 * CHEB d = cheb_a (dinv (b - A u) + cheb_c d), y = u + d (d not read when cheb_c == 0);
 * DOT y = A u and partial[t] = sum of u . A u over tile t (FP64, length vh_matvec_tiles());
 *     y may be NULL in a DOT pass (only the partials are kept), and in a RESID pass that
 *     writes y32 (only the FP32 copy is kept);
 * CHEB3 y = u + cheb_a (dinv (b - A u) + cheb_c (u - y)), y holding the previous iterate on
 * entry (or prev holding it; not read when cheb_c == 0): the three-term form of CHEB without a
 * direction array; CHEB3D the same step and the two dots of the field list below.
 * storage = VH_F32 with dtype = VH_F64: u in FP32 for a DOT pass with the AXPY load (the
 * mixed-precision Krylov loop reads its FP32 preconditioned residual directly); the load-time
 * combination's w is in the compute dtype.
 * (Load-time combination and extra outputs of DOT / RESID: see the w, sdev, u_out, y32 fields.)
 * JACOBI/CHEB/CHEB3 with dinv == NULL compute D^-1 = 1 / (kdiag[c] * sum of the incident E)
 * at free DOFs and 1 at fixed ones (kdiag: the diagonal of ke per component, which an
 * axis-aligned brick has equal at its 8 corners).
 * Replaces: the smoothing sweep and residual of Multigrid.vcycle src/voxel_mg.py:166-174,
 * the Chebyshev step of donor_yang_gmg multigrid_v4.py:860-889, the matvec + p.q of fcg
 * src/voxel_mg.py:197-199 and of donor pcg simp_r2.py:258-262. */
typedef struct vh_epilogue {
  int mode;
  int fixed_zero;
  const void* b;
  const void* dinv;
  void* d;
  double* partial;
  double omega, cheb_a, cheb_c;
  int lx;   /* x planes per workgroup chunk, 0 = automatic */
  int tile; /* (y, z) tile variant: 0 = automatic, 1 = 16 x 32 nodes, 2 = 8 x 32 nodes, 3 = 4 x 8 nodes (one
             * wave), 4 = wide (one workgroup per CU: 22 x 32 nodes in FP64, 30 x 32 in FP32) */
  double kdiag[3];
  int flags; /* VH_PREV_ZERO: CHEB3's previous iterate is zero (y is then not read);
             * VH_W_F32: w is stored in FP32 (FP64 RESID pass with the load-time combination only:
             * the mixed-precision Krylov loop forms its search direction in FP32) */
  /* Load-time combination (voxel-cube blocks), the combined input written to u_out at every
   * DOF:  DOT / RESID read u + s w with s = sdev[sidx] (device FP64 scalar);  CHEB3 reads
   * load_scale * (u (.) w) (elementwise product: with u = b, w = dinv and load_scale = 1/theta
   * this is the first Chebyshev step from a zero iterate, so the pass applies steps 1 and 2; the
   * previous iterate is zero: VH_PREV_ZERO or cheb_c = 0 is required, and it is not read).
   * RESID also writes a float32 copy of y to y32 and, with partial != NULL, per-tile sums of y^2. */
  const void* w;
  const double* sdev;
  int sidx;
  void* u_out;
  float* y32;
  double load_scale;
  /* CHEB3 / CHEB3D: the previous iterate, when it is not held by y on entry (y is then written
   * only).  CHEB3D (the last smoothing step of a V-cycle fused with the Polak-Ribiere dots of the
   * Krylov loop): partial[t] = sum of b . y and partial[n_tiles + t] = sum of b . (y - zold) over
   * the DOFs of tile t (FP64; length 2 vh_matvec_tiles()). */
  const void* prev;
  const void* zold;
} vh_epilogue;
#define VH_PREV_ZERO 1
#define VH_W_F32 2
int vh_matvec_fused(int dtype, int storage, const double* blocks72, const void* E, const void* u, void* y,
                    int nx, int ny, int nz, int mask_kind, const uint8_t* mask, const vh_epilogue* ep,
                    void* stream);
/* The DOT pass of vh_matvec_fused (y = A u, or with y == NULL only the dot) with no workspace and
 * no allocation: *pq (one FP64 device value) = u . A u.  The per-tile partial sums go to an
 * array that the library holds in device memory for the life of the process (512 KiB per
 * GPU, 65,536 tiles: grids whose default plan needs more tiles are rejected with -2).  The first
 * call on a device looks up that device's copy of the array (no allocation then either).
 * Calls that overlap on one GPU (two streams, two threads) race on the array: serialise them.
 * No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). */
int vh_matvec_dot_noalloc(int dtype, const double* blocks72, const void* E, const void* u, void* y, int nx, int ny,
                          int nz, int mask_kind, const uint8_t* mask, double* pq, void* stream);
/* 1 if the 72 block values have the voxel-cube pattern (8 distinct values; the form the
 * load-time combinations and the batched product require), 0 otherwise.  A query of the
 * kernel's coefficient form (host), with no counterpart in the reference. */
int vh_blocks_are_cube(const double* blocks72);
/* Number of tiles (= DOT partials) and workgroups the product launches for these arguments
 * (blocks72 host pointer or NULL for the voxel-cube pattern; lx/tile as in vh_epilogue). */
int vh_matvec_tiles(int dtype, int storage, const double* blocks72, int nx, int ny, int nz, int lx, int tile,
                    int* n_tiles, int* n_workgroups);

/* Node-owned product with the dense 24x24 ke (host pointer): the dense8 / node control.  The
 * library keeps a device copy per distinct ke (per device and dtype, for the life of the process),
 * so repeated calls with a matrix seen before copy nothing and may be captured in a HIP graph; a
 * new ke inside a capture is an error.
 * Replaces: dense8 code/hex_modal_cuda/cuda.py:47-79 and node_matvec_fp64
 * experiments/session_a/donor/src/gpu_fem/cuda_operators.py:204-285. */
int vh_dense_product(int dtype, const double* ke576, const void* E, const void* u, void* y, int nx, int ny, int nz,
                     void* stream);

/* Element-per-lanes layout control: each element worked by 24 lanes (node x component) of a
 * 32-lane half of a wave64, transforms by lane shuffles, y = K(E) u (unconstrained; y is zeroed
 * first and the contributions are added with atomics, so the summation order is not fixed).
 * Replaces: modal32 code/hex_modal_layout_v2/variants.py:95-117. */
int vh_lanes24_product(int dtype, const double* blocks72, const void* E, const void* u, void* y, int nx, int ny,
                       int nz, void* stream);

/* Jacobi diagonal (invert = 0) or inverse diagonal with 1 at fixed DOFs (invert = 1).
 * Replaces: jacobi_diag_fp64 cuda_operators.py:371-421 / OperatorSuite.diagonal :763-792;
 * NumpyProduct.diagonal src/voxel_mg.py:55-58 and Level.inv_diag :126-127. */
int vh_diagonal(int dtype, const double* ke576, const void* E, void* out, int nx, int ny, int nz, int mask_kind,
                const uint8_t* mask, int invert, void* stream);

/* Element energy ce[e] = u_e^T ke u_e of the masked u; if dc != NULL also
 * dc[e] = -penal * dE * rho[e]^(penal-1) * ce[e].
 * Replaces: elem_compliance_fp64 cuda_operators.py:304-355 and the sensitivity of
 * donor_yang_gmg solver_v4.py:246-251. */
int vh_element_energy(int dtype, const double* blocks72, const void* u, void* ce, int nx, int ny, int nz,
                      int mask_kind, const uint8_t* mask, const void* rho, void* dc, double penal, double dE,
                      void* stream);

/* Restriction rc = Mc P^T (Mf rf) from the fine grid (nx, ny, nz even) to its 2x coarsening.
 * Replaces: restrict src/voxel_mg.py:83-89,100-105 with the masks of :169-170. */
int vh_restrict(int dtype, const void* rf, void* rc, int nx, int ny, int nz, int fmask_kind, const uint8_t* fmask,
                int cmask_kind, const uint8_t* cmask, void* stream);
/* Prolongation onto the fine grid: xf += Mf P ec (assign = 0) or xf = Mf P ec (assign = 1).
 * Replaces: prolong src/voxel_mg.py:74-80,92-97 and the correction of :172. */
int vh_prolong(int dtype, const void* ec, void* xf, int nx, int ny, int nz, int mask_kind, const uint8_t* mask,
               int assign, void* stream);
/* Ec = 2x2x2 mean of Ef.  Replaces: coarsen_modulus src/voxel_mg.py:108-111. */
int vh_coarsen_modulus(int dtype, const void* Ef, void* Ec, int nx, int ny, int nz, void* stream);

/* Dense constrained operator A = M K M + (I - M) of a small grid (FP64, device arrays: ke the
 * 24x24 element matrix in its native node order, E the element moduli, A row-major n x n with
 * row stride ld >= n = 3 (nx+1)(ny+1)(nz+1); every entry is written).  K_ij sums E_e ke over the
 * shared elements in ascending element order.
 * Replaces: assemble_dense src/voxel_mg.py:61-70 (host scipy assembly of the coarsest operator,
 * re-run at every density update). */
int vh_assemble_dense(const double* ke, const double* E, int nx, int ny, int nz, int mask_kind, const uint8_t* mask,
                      double* A, int ld, void* stream);
/* Galerkin coarse operator M_c (P^T A_f P) M_c + (I - M_c) from 81 probe products (FP64 device
 * arrays; (nx, ny, nz) the COARSE element grid, n its DOF count): row p = ((px*3+py)*3+pz)*3+c of
 * G (81 x n) holds R A_f P v_p for v_p = the sum of the unit vectors e_(J, c) over the coarse
 * nodes J with (Jx, Jy, Jz) = (px, py, pz) mod 3.  Writes the entries of each row's 27-node
 * neighbourhood; A (row stride ld >= n) must be zero on entry (all other entries are 0).
 * Replaces: the Galerkin triple product assemble_level1_galerkin multigrid_v4.py:248-270,747-772
 * (host sparse products). */
int vh_galerkin_scatter(const double* G, int nx, int ny, int nz, int mask_kind, const uint8_t* mask, double* A, int ld,
                        void* stream);

/* Galerkin coarse operator of a non-coarsest level as a symmetric 27-point node-block stencil
 * (dtype F64 or F32, device arrays; (nx, ny, nz) the COARSE element grid, nn its node count).
 * G is the 81 x 3nn probe matrix of vh_galerkin_scatter (FP64: R A_f P v_p, A_f the finer
 * level's constrained operator, R and P unmasked).  U (14 x 9 x nn) receives the blocks
 * K[I, I + d] of the upper half of the offsets d in {-1,0,1}^3 (offset index
 * k = ((dx+1)*3+(dy+1))*3+(dz+1) >= 13, slot k - 13), coefficient-major:
 * U[((slot*3 + r)*3 + c)*nn + I] = K[3I + r, 3(I + d) + c], 0 where I + d is outside the grid.
 * K is the unmasked P^T A_f P; the coarse mask is applied by vh_stencil_apply.
 * Replaces: the Galerkin coarse operators of donor_yang_gmg multigrid_v4.py:132-270 (level 1,
 * CSR from element triple products) and :747-772 (deeper levels, sparse R K P). */
int vh_stencil_scatter(int dtype, const double* G, void* U, int nx, int ny, int nz, void* stream);
/* y = M K M u + (I - M) u with the stencil K of vh_stencil_scatter (one thread per node), and
 * the epilogues of vh_matvec_fused's RESID (with y32), JACOBI and CHEB3 (prev, VH_PREV_ZERO)
 * passes (fixed_zero as there).  The smoothing passes take dinv (kdiag is not used); ep's
 * d, partial, w, zold and u_out must be NULL.
 * Replaces: the CSR SpMV of the coarse Chebyshev smoothing and residuals of donor_yang_gmg
 * multigrid_v4.py:934-1008. */
int vh_stencil_apply(int dtype, const void* U, const void* u, void* y, int nx, int ny, int nz, int mask_kind,
                     const uint8_t* mask, const vh_epilogue* ep, void* stream);

/* y = mask * (A r) for a dense row-major n x n matrix with leading dimension ld (ld >= n,
 * ld % 4 == 0, A 16-byte aligned; the padding columns are not read): the coarsest solve with the
 * explicit inverse of the coarsest operator.  a_dtype is the matrix type, v_dtype the type of r
 * and y (an FP64 matrix may take FP32 vectors: the coarsest solve in FP64 inside an FP32
 * V-cycle; accumulation in the matrix type).
 * Replaces: Multigrid.coarse_solve src/voxel_mg.py:152-159. */
int vh_dense_mv(int a_dtype, int v_dtype, int n, int ld, const void* A, const void* r, void* y, const uint8_t* mask,
                void* stream);

/* Deterministic two-stage reductions with device-resident FP64 scalars.
 * vh_reduce writes FP64 partials (count = 2 for DOT2/DOTDIFF, laid out [j][part]) of length
 * vh_reduce_parts(n); vh_finish sums them into scal[dst..] and applies a scalar op.
 * Replaces: the cp.dot / cp.linalg.norm host reductions of fcg src/voxel_mg.py:187-211 and
 * donor pcg simp_r2.py:242-285. */
int vh_reduce_parts(size_t n);
int vh_reduce(int dtype, int kind, size_t n, const void* x, const void* y, const void* w, double* partial,
              void* stream);
int vh_finish(const double* partial, int nparts, int count, double* scal, int dst, int op, void* stream);
/* x += alpha p; if update_r: r -= alpha q and, with dinv, z = dinv r and partial r.z
 * (alpha = scal[VH_S_ALPHA]).  Replaces: the update lines of donor pcg simp_r2.py:266-284. */
int vh_cg_update(int dtype, size_t n, void* x, const void* p, void* r, const void* q, const void* dinv, void* z,
                 const double* scal, int update_r, double* partial, void* stream);
/* znew = zin converted to dtype_v, with FP64 partials [r.znew | r.(znew - zold)] (layout of
 * VH_RED_DOTDIFF, vh_reduce_parts(n) each); r, zold, znew are of dtype_v (FP64 with an FP32 or
 * FP64 input, or all FP32): the V-cycle output turned into the next preconditioned vector of a
 * flexible CG step together with both Polak-Ribiere dots.
 * Replaces: rz_new = r @ z_new and beta = r @ (z_new - z) / rz of fcg src/voxel_mg.py:209-211. */
int vh_cvt_dotdiff(int dtype_in, int dtype_v, size_t n, const void* zin, const void* r, const void* zold, void* znew,
                   double* partial, void* stream);
/* z = dinv r and partial r.z.  Replaces: z = M_inv * r; rz = r.z of simp_r2.py:239-242,282-283. */
int vh_jacobi_dot(int dtype, size_t n, const void* r, const void* dinv, void* z, double* partial, void* stream);
/* p = z + beta p (beta = scal[VH_S_BETA]); if zcopy != NULL also zcopy = z in the same pass.
 * Replaces: p = z + beta p of fcg src/voxel_mg.py:212-213 / donor pcg simp_r2.py:286. */
int vh_xpby(int dtype, size_t n, const void* z, void* p, const double* scal, void* zcopy, void* stream);
/* y = a x (w) + b y with host scalars (w optional, elementwise). */
int vh_axpby(int dtype, size_t n, double a, const void* x, const void* w, double b, void* y, void* stream);
/* y = x * s (accumulate: y += x * s), s = scale or *dscale (device), inverted if inv_scale;
 * converts between float64 and float32, or float32 to bfloat16 (dtype_out VH_BF16: rounded to
 * nearest even, no accumulate; the BF16 storage of a V-cycle iterate).  Replaces: the FP32 cast
 * of the scaled residual and the FP64 update of donor pcg_ir simp_r2.py:372-379, and the BF16
 * cast of the fine smoother's input in donor_yang_gmg multigrid_v4.py:851-856. */
int vh_convert(int dtype_in, int dtype_out, size_t n, const void* x, void* y, double scale, const double* dscale,
               int inv_scale, int accumulate, void* stream);
/* y[i] = 2 u_i - 1 in [-1, 1) with u_i = (s_i >> 11) 2^-53, s_i the (i+1)-th output of SplitMix64 seeded
 * with `seed` (Steele, Lea, Flood, OOPSLA 2014).  No counterpart in the reference: the deterministic
 * start vector of the Chebyshev spectrum estimate, drawn on the device (the reference's seeded host
 * normals cost about 0.8 host seconds per 100 M values), reproducible bit for bit by integer arithmetic. */
int vh_fill_splitmix(int dtype, size_t n, unsigned long long seed, void* y, void* stream);

/* Cone filter y = post_div^-1 H (x / pre_div) with radius rmin (pre_div/post_div optional).
 * Replaces: cone_stencil_fp64 filter_r2.py:40-76, ConeFilter.forward/transpose :171-178. */
int vh_cone_filter(int dtype, const void* x, const void* pre_div, const void* post_div, void* y, int nx, int ny,
                   int nz, double rmin, void* stream);
/* OC probe: xnew = clip(x sqrt(max(-dc/dv/lmid, 0)), max(x-move,0), min(x+move,1)) and FP64
 * partials of sum xnew * w (nparts blocks).  Replaces: one bisection probe of oc_top88_gpu
 * experiments/session_h/run_w1b.py:82-94 (with w = H(1/s)/n the filtered volume). */
int vh_oc_probe(int dtype, size_t n, const void* x, const void* dc, const void* dv, const void* w, double lmid,
                double move, void* xnew, double* partial, int nparts, void* stream);
/* count (1 .. 16) OC bisection probes in one pass: partial[k * nparts + b] is block b's FP64 sum of
 * xnew(lmids[k]) * w, bitwise the partials vh_oc_probe returns for that multiplier (xnew is not
 * written).  Replaces: consecutive probes of the bisection of oc_top88_gpu
 * experiments/session_h/run_w1b.py:82-94 (its midpoints are known in advance along a branch of
 * the bisection, so several of them are evaluated in one read of the design). */
int vh_oc_probe_multi(int dtype, size_t n, const void* x, const void* dc, const void* dv, const void* w,
                      const double* lmids, int count, double move, double* partial, int nparts, void* stream);

/* Heaviside projection of the filtered density rho_bar (FP64, n elements): rho = (tanh(beta eta)
 * + tanh(beta (rho_bar - eta))) / (tanh(beta eta) + tanh(beta (1 - eta))) (identity for beta <=
 * 1e-12), its derivative drho, the modulus E = Emin + (E0 - Emin) rho^penal, and FP64 per-block
 * sums of rho (nparts blocks: the volume); every output optional.
 * Replaces: project / dproject simp_r2.py:388-404 and the modulus of run_simp :650-652. */
int vh_heaviside(size_t n, const double* rho_bar, double beta, double eta, double penal, double E0, double Emin,
                 double* rho, double* drho, double* E, double* partial, int nparts, void* stream);
/* Candidate design of one OC bisection probe: rho_new = clip(clip(rho sqrt(-min(dc, -1e-30) /
 * (lmid max(dv, 1e-30))), rho - move, rho + move), rho_lb, 1) (FP64).
 * Replaces: the design() closure of oc_update simp_r2.py:497-507. */
int vh_oc_design(size_t n, const double* rho, const double* dc, const double* dv, double lmid, double move,
                 double rho_lb, double* rho_new, void* stream);

/* ------------------------------------------------------------------ batched problems
 * Many voxel problems (grids of different sizes, several load cases each) run each pass of
 * their solver in ONE launch.  A batch handle holds a descriptor table in device memory,
 * built once (set-up; it allocates) and launched any number of times (also inside captured
 * HIP graphs); vh_batch_destroy releases it.  These entry points have no single counterpart
 * in the reference, which solves one problem per call (fcg src/voxel_mg.py:181-217 on one
 * grid); they batch the per-problem operations declared above. */
typedef struct vh_product_item {
  const double* blocks72;       /* host, voxel-cube pattern required */
  const void* E;
  const void* u;
  void* y;
  int nx, ny, nz;
  const uint8_t* mask;          /* VH_MASK_ARRAY only */
  vh_epilogue ep;               /* mode, operands and scalars of this problem's pass; ep.lx fixes the
                                   x chunk length (0: planned for the whole batch) */
} vh_product_item;
/* Products of all items in one launch (same dtype, mode, mask kind, tile and load-time
 * combination).  DOT/RESID partials: every item passes the same ep.partial base; item i writes
 * its tiles at [offs[i], offs[i+1]) of it (offs from vh_batch_offsets, for vh_seg_finish).
 * tiles_out (n ints, optional) receives each item's tile count.  tile: 2 = 8 x 32 nodes,
 * 5 = 16 x 16 nodes (batched products only), anything else 16 x 32.  Modes RAW, RESID, DOT,
 * JACOBI, CHEB3.  As in vh_matvec_fused, an FP64 RESID pass may take an FP32 w
 * (ep.flags VH_W_F32, every item alike) and may omit y when it writes ep.y32. */
int vh_product_batch_create(int dtype, int mask_kind, int tile, int n, const vh_product_item* items, void** handle,
                            int* tiles_out);

/* Segments of a concatenated vector (item i owns elements [sum_{j<i} len_j, ... + len_i)). */
int vh_segments_create(int n, const long long* lengths, void** handle);
/* Per-segment |x|^2 (VH_RED_NORM2) or x.y (VH_RED_DOT): one FP64 partial per workgroup. */
int vh_seg_reduce(void* segments, int dtype, int kind, const void* x, const void* y, double* partial, void* stream);
/* Per-segment cvt_dotdiff: znew = zin (converted), partials [r.znew | r.(znew - zold)]. */
int vh_seg_dotdiff(void* segments, int dtype_in, int dtype_v, const void* zin, const void* r, const void* zold,
                   void* znew, double* partial, void* stream);
/* One workgroup per item: sums partial[j * total + offs[i] .. offs[i+1]) (j < count) into
 * scal[i * stride + dst ..] and runs the scalar recurrence `op` on that item's scalar block.
 * offs: a device prefix table from vh_batch_offsets (product tiles or segment blocks). */
int vh_seg_finish(const double* partial, int total, const int* offs, int n, int count, double* scal, int stride,
                  int dst, int op, void* stream);

typedef struct vh_transfer_item {
  const void* src;              /* restrict: fine residual; prolong: coarse correction */
  void* dst;                    /* restrict: coarse rhs; prolong: fine iterate */
  int nx, ny, nz;               /* FINE element grid (even extents) */
  int fmask_kind, cmask_kind;
  const uint8_t* fmask;
  const uint8_t* cmask;
} vh_transfer_item;
/* Restriction (restrict_op = 1) or prolongation (0; assign = 1 overwrites, else adds) of all
 * items in one launch, as vh_restrict / vh_prolong. */
int vh_transfer_batch_create(int dtype, int restrict_op, int assign, int n, const vh_transfer_item* items,
                             void** handle);

typedef struct vh_dense_item {
  const void* A;                /* row-major, leading dimension ld (ld % 4 == 0, 16-byte aligned) */
  int ld, n;
  const void* r;
  void* y;
  const uint8_t* mask;
} vh_dense_item;
/* The coarsest solves y = mask * (A r) of all items in one launch, as vh_dense_mv.  Consecutive
 * items with the same A, n, ld and mask (the load cases of one problem) are applied in one pass
 * over A, up to 4 at a time; each result is bitwise that of a separate pass. */
int vh_dense_batch_create(int a_dtype, int v_dtype, int n, const vh_dense_item* items, void** handle);

/* Launch a product, transfer or dense batch on `stream`. */
int vh_batch_launch(void* handle, void* stream);
/* Device prefix table (n + 1 ints: tiles, segment blocks, lines or row blocks) of a handle. */
int vh_batch_offsets(void* handle, const int** offs, int* n, int* total);
int vh_batch_destroy(void* handle);

#ifdef __cplusplus
}
#endif
#endif
