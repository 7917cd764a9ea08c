"""CPU oracle for the Hex8 voxel element algebra and the matrix-free product.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this module; only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the checker
and as the timed CPU baseline.

Restates, in NumPy:
  * the fully integrated 2x2x2-Gauss Hex8 stiffness      code/hex_modal/modal.py:19-40
  * the parity-block (modal) factorisation C = T K T^T/64  code/hex_modal/modal.py:9-16,55-77
  * the kernel lane map, butterfly and lane emulation     src/voxel_kernels.py:46-97
  * the assembled element-loop product                    src/voxel_kernels.py:100-110
  * the author-layout DOF table and the CPU product       src/voxel_mg.py:27-70
  * the natural <-> author layout transposes              experiments/session_a/dependencies/common.py:35-46
  * the natural-layout product/energy/diagonal of the C++ oracle
                                                          experiments/session_a/inputs/reference.cc:16-39

Layouts (author): element e = (ex*ny + ey)*nz + ez, node n = (ix*(ny+1) + iy)*(nz+1) + iz,
DOF 3n+c (x slowest, z fastest, components interleaved).

Parity status (what pins each function; see tests/test_oracle_*.py):
  quadrature        closed form: ke_unit_3d.npy fixture, symmetry, 6 rigid-body null modes,
                    fixture ke == unit/nx; C++ element() of oracle/_ref where built
  Modal             T T^T = 8I, reconstruction T^T C T = K, off-block share < 2e-14,
                    rejection of a non-brick ke
  emulate / product brute force against assembled_reference and against dense assembly;
                    oracle/_ref product(); rigid-body modes in the null space
  NumpyProduct      same as above, diagonal == diag(assemble_dense)
"""
from __future__ import annotations

import numpy as np

# Native Hex8 node order of `ke` (src/voxel_mg.py:27-28, code/hex_modal/modal.py:9-10).
XYZ = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
# binary lane number of each native node: lane = x + 2y + 4z
NATIVE_LANE = XYZ[:, 0] + 2 * XYZ[:, 1] + 4 * XYZ[:, 2]
# native node of each binary lane (inverse permutation)
LANE_NATIVE = np.argsort(NATIVE_LANE)


def hadamard8():
    """H[p, l] = (-1)^popcount(p & l) over binary lane numbers l."""
    p = np.arange(8)[:, None]
    l = np.arange(8)[None, :]
    bits = np.vectorize(lambda v: bin(int(v)).count("1"))(p & l)
    return np.where(bits % 2 == 0, 1.0, -1.0)


H8 = hadamard8()
# T maps the native-ordered 24-vector of an element to its modal coefficients:
# row 3p+c, column 3n+c carries H[p, lane(n)].
T24 = np.kron(H8[:, NATIVE_LANE], np.eye(3))
# Parity block g couples (mode g^1, comp x), (mode g^2, comp y), (mode g^4, comp z).
BLOCK_INDEX = np.array([[3 * (g ^ 1), 3 * (g ^ 2) + 1, 3 * (g ^ 4) + 2] for g in range(8)])


def quadrature(lengths, nu):
    """24x24 stiffness of an axis-aligned brick with unit Young's modulus.

    2x2x2 Gauss integration of B^T D B with engineering-strain Voigt order
    (xx, yy, zz, xy, yz, zx); nodes in the native order XYZ (code/hex_modal/modal.py:19-40).
    """
    h = np.asarray(lengths, dtype=np.float64)
    lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = 0.5 / (1.0 + nu)
    D = np.diag([lam + 2 * mu] * 3 + [mu] * 3)
    D[:3, :3] += lam * (1.0 - np.eye(3))
    s = 2.0 * XYZ - 1.0                                  # node sign per axis
    g = 1.0 / np.sqrt(3.0)
    ke = np.zeros((24, 24))
    for q in s * g:                                      # the 8 Gauss points
        # dN_i/dx_a = s_ia / (4 h_a) * prod_{b != a} (1 + s_ib q_b)
        dN = np.empty((8, 3))
        for a in range(3):
            other = [b for b in range(3) if b != a]
            dN[:, a] = s[:, a] / (4.0 * h[a]) * np.prod(1.0 + s[:, other] * q[other], axis=1)
        B = np.zeros((6, 24))
        for i in range(8):
            bx, by, bz = dN[i]
            B[0, 3 * i] = bx
            B[1, 3 * i + 1] = by
            B[2, 3 * i + 2] = bz
            B[3, 3 * i], B[3, 3 * i + 1] = by, bx
            B[4, 3 * i + 1], B[4, 3 * i + 2] = bz, by
            B[5, 3 * i], B[5, 3 * i + 2] = bz, bx
        ke += (B.T @ D @ B) * (np.prod(h) / 8.0)
    return ke


def butterfly(values):
    """Unnormalised 8-point Walsh-Hadamard transform along axis 1 in binary lane order
    (code/hex_modal/modal.py:43-52): out[:, p] = sum_l H[p, l] v[:, l]."""
    v = np.array(values, dtype=np.float64, copy=True)
    for stride in (1, 2, 4):
        idx = np.arange(8)
        lo = idx[(idx & stride) == 0]
        a, b = v[:, lo].copy(), v[:, lo + stride].copy()
        v[:, lo], v[:, lo + stride] = a + b, a - b
    return v


class Modal:
    """Parity-block factorisation of an element matrix (code/hex_modal/modal.py:55-77).

    full = T K T^T / 64; blocks[g] = full[BLOCK_INDEX[g]][:, BLOCK_INDEX[g]].  Raises ValueError
    when the Frobenius share outside the blocks exceeds `tolerance` (the element is not an
    axis-aligned isotropic brick).
    """

    def __init__(self, ke, tolerance=2e-14):
        self.ke = np.asarray(ke, dtype=np.float64).reshape(24, 24)
        full = T24 @ self.ke @ T24.T / 64.0
        kept = np.zeros_like(full)
        self.blocks = np.empty((8, 3, 3))
        for g in range(8):
            ix = np.ix_(BLOCK_INDEX[g], BLOCK_INDEX[g])
            self.blocks[g] = full[ix]
            kept[ix] = full[ix]
        self.discarded_relative = float(np.linalg.norm(full - kept) / np.linalg.norm(full))
        if not np.isfinite(self.discarded_relative) or self.discarded_relative > tolerance:
            raise ValueError(f"Element is outside the declared parity-block model: "
                             f"{self.discarded_relative}")
        self.full_modal = full
        self.sparse_modal = kept
        self.reconstructed = T24.T @ kept @ T24

    def apply(self, values):
        """K u_e for a batch of native-ordered element vectors (n, 24) via the blocks."""
        v = np.asarray(values, dtype=np.float64).reshape(-1, 8, 3)
        m = butterfly(v[:, LANE_NATIVE]).reshape(-1, 24)     # modal index 3p+c
        out = np.zeros_like(m)
        for g in range(8):
            out[:, BLOCK_INDEX[g]] = m[:, BLOCK_INDEX[g]] @ self.blocks[g].T
        o = butterfly(out.reshape(-1, 8, 3))
        return o[:, NATIVE_LANE].reshape(-1, 24)


# ------------------------------------------------------------------------- grid indexing
def n_nodes(shape):
    nx, ny, nz = (int(s) for s in shape)
    return (nx + 1) * (ny + 1) * (nz + 1)


def element_nodes(shape):
    """(n_el, 8) global node of every (element, binary lane) (src/voxel_kernels.py:46-53)."""
    nx, ny, nz = (int(s) for s in shape)
    e = np.arange(nx * ny * nz)
    ez, ey, ex = e % nz, (e // nz) % ny, e // (nz * ny)
    lane = np.arange(8)
    lx, ly, lz = lane & 1, (lane >> 1) & 1, (lane >> 2) & 1
    return ((ex[:, None] + lx) * (ny + 1) + ey[:, None] + ly) * (nz + 1) + ez[:, None] + lz


def edof_author(shape):
    """(n_el, 24) author-layout DOFs in the native node order of ke (src/voxel_mg.py:32-38)."""
    nodes = element_nodes(shape)[:, NATIVE_LANE]          # native order
    return (3 * nodes[:, :, None] + np.arange(3)).reshape(-1, 24)


def lane_transform(v):
    """Kernel butterfly on (n, 8) lane data: for bit in 1,2,4: lane&bit ? other - v : v + other
    (src/voxel_kernels.py:56-62; code/hex_modal_cuda/cuda.py:9-16)."""
    v = np.array(v, copy=True)
    lane = np.arange(8)
    for bit in (1, 2, 4):
        other = v[:, lane ^ bit]
        v = np.where((lane & bit) != 0, other - v, v + other)
    return v


def emulate(path, shape, ke, E, u, dtype=np.float64):
    """Lane-level replica of the modal8 / dense8 kernels at `dtype` (src/voxel_kernels.py:65-97).

    modal8: transform each component, apply the block entries with the cross-component
    partners lane^3 (x<->y), lane^5 (x<->z), lane^6 (y<->z), transform back, scale by E,
    scatter.  dense8: lane L (native node L^((L&2)>>1)) accumulates row 3*native+c of ke
    against every lane's u.
    """
    nodes = element_nodes(shape)
    nn = n_nodes(shape)
    E = np.asarray(E, dtype=dtype)
    u = np.asarray(u, dtype=dtype)
    vx, vy, vz = (u[3 * nodes + c] for c in range(3))
    lane = np.arange(8)
    if path == "modal8":
        C = np.asarray(Modal(ke).blocks, dtype=np.float64).reshape(72).astype(dtype)
        vx, vy, vz = lane_transform(vx), lane_transform(vy), lane_transform(vz)
        gx, gy, gz = 9 * (lane ^ 1), 9 * (lane ^ 2), 9 * (lane ^ 4)
        fx = C[gx] * vx + C[gx + 1] * vy[:, lane ^ 3] + C[gx + 2] * vz[:, lane ^ 5]
        fy = C[gy + 3] * vx[:, lane ^ 3] + C[gy + 4] * vy + C[gy + 5] * vz[:, lane ^ 6]
        fz = C[gz + 6] * vx[:, lane ^ 5] + C[gz + 7] * vy[:, lane ^ 6] + C[gz + 8] * vz
        fx, fy, fz = lane_transform(fx), lane_transform(fy), lane_transform(fz)
    elif path == "dense8":
        K = np.asarray(ke, dtype=np.float64).reshape(576).astype(dtype)
        native = lane ^ ((lane & 2) >> 1)
        fx, fy, fz = np.zeros_like(vx), np.zeros_like(vx), np.zeros_like(vx)
        for j in range(8):
            src = j ^ ((j & 2) >> 1)
            ax, ay, az = vx[:, [src]], vy[:, [src]], vz[:, [src]]
            base = 72 * native + 3 * j
            fx = fx + K[base] * ax + K[base + 1] * ay + K[base + 2] * az
            fy = fy + K[base + 24] * ax + K[base + 25] * ay + K[base + 26] * az
            fz = fz + K[base + 48] * ax + K[base + 49] * ay + K[base + 50] * az
    else:
        raise ValueError(path)
    y = np.zeros(3 * nn, dtype=dtype)
    for c, f in enumerate((fx, fy, fz)):
        np.add.at(y, 3 * nodes + c, (E[:, None] * f).astype(dtype))
    return y


def assembled_reference(shape, ke, E, u):
    """Element-by-element y[dofs] += E_e ke u[dofs] in FP64 (src/voxel_kernels.py:100-110)."""
    edof = edof_author(shape)
    y = np.zeros(3 * n_nodes(shape))
    for e in range(edof.shape[0]):
        d = edof[e]
        y[d] += E[e] * (ke @ u[d])
    return y


class NumpyProduct:
    """CPU product y = K(E) u with an np.add.at scatter (src/voxel_mg.py:41-58)."""

    def __init__(self, shape, ke, dtype=np.float64):
        self.shape, self.dtype = tuple(int(s) for s in shape), dtype
        self.ke = np.asarray(ke, dtype=dtype)
        self.edof = edof_author(self.shape)
        self.ndof = 3 * n_nodes(self.shape)

    def raw(self, u, E):
        y = np.zeros(self.ndof, dtype=self.dtype)
        local = u[self.edof] @ self.ke.T
        np.add.at(y, self.edof, (E[:, None] * local).astype(self.dtype))
        return y

    def diagonal(self, E):
        d = np.zeros(self.ndof, dtype=self.dtype)
        np.add.at(d, self.edof, (E[:, None] * np.diag(self.ke)[None, :]).astype(self.dtype))
        return d


def assemble_dense(shape, ke, E, mask):
    """Dense M K M + (I - M) in FP64 (src/voxel_mg.py:61-70)."""
    edof = edof_author(shape)
    n = 3 * n_nodes(shape)
    K = np.zeros((n, n))
    for e in range(edof.shape[0]):
        K[np.ix_(edof[e], edof[e])] += E[e] * ke
    m = np.asarray(mask, dtype=np.float64)
    return m[:, None] * K * m[None, :] + np.diag(1.0 - m)


def element_energy(shape, ke, u):
    """c_e = u_e^T ke u_e per element (donor cuda_operators.py:304-355, solver_v4.py:246-251)."""
    edof = edof_author(shape)
    ue = u[edof]
    return np.einsum("ei,ij,ej->e", ue, ke, ue)


# ------------------------------------------------------------------------- layouts
def to_author(a, shape, nodal=False):
    """Natural (z, y, x[, 3]) -> author (x, y, z[, 3]) (common.py:35-39)."""
    x, y, z = (int(s) for s in shape)
    if nodal:
        return np.ascontiguousarray(np.asarray(a).reshape(z + 1, y + 1, x + 1, 3)
                                    .transpose(2, 1, 0, 3)).ravel()
    return np.ascontiguousarray(np.asarray(a).reshape(z, y, x).transpose(2, 1, 0)).ravel()


def to_natural(a, shape, nodal=False):
    """Author (x, y, z[, 3]) -> natural (z, y, x[, 3]) (common.py:42-46)."""
    x, y, z = (int(s) for s in shape)
    if nodal:
        return np.ascontiguousarray(np.asarray(a).reshape(x + 1, y + 1, z + 1, 3)
                                    .transpose(2, 1, 0, 3)).ravel()
    return np.ascontiguousarray(np.asarray(a).reshape(x, y, z).transpose(2, 1, 0)).ravel()


def natural_product(shape, rho, u, ke, mask, emin, emax, penal):
    """Natural-layout (y, ce, diag) exactly as the C++ oracle defines them
    (experiments/session_a/inputs/reference.cc:16-39): u masked on read, identity rows and
    unit diagonal at fixed DOFs, ce = u_e^T ke u_e of the masked u."""
    x, y_, z = (int(s) for s in shape)
    E = emin + (emax - emin) * np.asarray(rho) ** penal
    Ea = to_author(E, shape)
    ua = to_author(np.where(mask != 0, u, 0.0), shape, nodal=True)
    prod = NumpyProduct(shape, ke)
    ya = prod.raw(ua, Ea)
    da = prod.diagonal(Ea)
    ce = to_natural(element_energy(shape, ke, ua), shape)
    yn = to_natural(ya, shape, nodal=True)
    dn = to_natural(da, shape, nodal=True)
    fixed = np.asarray(mask) == 0
    yn[fixed] = np.asarray(u)[fixed]
    dn[fixed] = 1.0
    return yn, ce, dn
