"""GPU parity of the operator kernels against the oracle, through the C-ABI.

Tolerances: FP64 relative L2 <= 1e-12 (the parity bar of this operator; the reference's own
gate is 1e-11, experiments/session_a/protocol.json:127-133); FP32 <= 1e-5 (same gate).
Index/mask work (fixed DOFs) is checked exactly.
"""
import warnings

import numpy as np
import pytest
import torch

from oracle import hexref as hx
from oracle import mgref as mg_ref

pytestmark = pytest.mark.gpu

KE = hx.quadrature([1, 1, 1], 0.3)


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else a
    return float(np.linalg.norm(np.asarray(a, dtype=np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


def case(shape, seed=7):
    rng = np.random.default_rng(seed)
    E = 1e-6 + (1 - 1e-6) * rng.random(int(np.prod(shape))) ** 3
    u = rng.standard_normal(3 * hx.n_nodes(shape))
    return E, u


SHAPES = [(3, 2, 4), (1, 1, 1), (2, 3, 5), (8, 4, 4), (17, 33, 40), (40, 20, 70), (5, 16, 32), (6, 17, 33),
          (9, 48, 96)]


@pytest.mark.parametrize("shape", SHAPES)
def test_parity_product_fp64_matches_the_oracle(shape):
    import voxel_kernels as vk
    print("case", shape, flush=True)
    E, u = case(shape)
    ref = hx.NumpyProduct(shape, KE).raw(u, E)
    y = vk.emulate("modal8", shape, KE, E, u, np.float64)
    assert rel(y, ref) < 1e-12
    if int(np.prod(shape)) < 2000:
        assert rel(y, hx.assembled_reference(shape, KE, E, u)) < 1e-12


@pytest.mark.parametrize("lx", [1, 2, 3, 7, 100])
def test_parity_product_is_independent_of_the_x_chunking(lx):
    from voxelhip import ops
    shape = (13, 18, 35)
    E, u = case(shape, 3)
    ref = hx.NumpyProduct(shape, KE).raw(u, E)
    y = torch.empty(u.size, dtype=torch.float64, device="cuda")
    ops.matvec(hx.Modal(KE).blocks, dev(E), dev(u), y, shape, lx=lx)
    assert rel(y, ref) < 1e-12


def test_parity_product_is_bitwise_deterministic():
    from voxelhip import ops
    shape = (24, 20, 36)
    E, u = case(shape, 5)
    Ed, ud = dev(E), dev(u)
    b = hx.Modal(KE).blocks
    y1, y2 = torch.empty_like(ud), torch.empty_like(ud)
    ops.matvec(b, Ed, ud, y1, shape)
    ops.matvec(b, Ed, ud, y2, shape)
    assert torch.equal(y1, y2)


@pytest.mark.parametrize("shape", [(3, 2, 4), (17, 33, 40)])
def test_parity_product_fp32(shape):
    import voxel_kernels as vk
    E, u = case(shape)
    ref = hx.NumpyProduct(shape, KE).raw(u, E)
    y = vk.emulate("modal8", shape, KE, E, u, np.float32)
    assert y.dtype == np.float32 and rel(y, ref) < 1e-5


def test_bf16_storage_product():
    from voxelhip import _lib as L, ops
    shape = (10, 12, 34)
    E, u = case(shape)
    ub = dev(u, torch.float32).to(torch.bfloat16)
    ref = hx.NumpyProduct(shape, KE).raw(ub.float().cpu().numpy().astype(np.float64), E)
    y = torch.empty(u.size, dtype=torch.float32, device="cuda")
    ops.matvec(hx.Modal(KE).blocks, dev(E, torch.float32), ub, y, shape, storage=L.BF16)
    assert rel(y, ref) < 1e-5          # exact bf16 inputs, FP32 arithmetic


@pytest.mark.parametrize("shape", [(3, 2, 4), (17, 33, 40)])
def test_dense_control_product(shape):
    import voxel_kernels as vk
    E, u = case(shape)
    ref = hx.NumpyProduct(shape, KE).raw(u, E)
    assert rel(vk.emulate("dense8", shape, KE, E, u, np.float64), ref) < 1e-12
    assert rel(vk.assembled_reference(shape, KE, E, u), ref) < 1e-12


def _masks(shape, rng):
    nn = hx.n_nodes(shape)
    x0 = np.ones((shape[0] + 1, nn // (shape[0] + 1), 3)); x0[0] = 0
    return {"x0": x0.ravel(), "array": (rng.random(3 * nn) > 0.15).astype(float)}


@pytest.mark.parametrize("kind", ["x0", "array"])
@pytest.mark.parametrize("shape", [(6, 5, 7), (18, 17, 34)])
def test_constrained_operator_and_epilogues(kind, shape):
    from voxelhip import _lib as L, ops
    rng = np.random.default_rng(11)
    E, u = case(shape, 4)
    m = _masks(shape, rng)[kind]
    prod = hx.NumpyProduct(shape, KE)
    lvl = mg_ref.Level(shape, prod, E, m, np.float64)
    Au = lvl.apply(u)
    b = rng.standard_normal(u.size)
    dinv = lvl.inv_diag
    d_old = rng.standard_normal(u.size)
    mask = ops.Mask.from_values(m, shape)
    assert mask.kind == (L.MASK_X0 if kind == "x0" else L.MASK_ARRAY)
    blocks = hx.Modal(KE).blocks
    Ed, ud, bd, dd = dev(E), dev(u), dev(b), dev(dinv)
    y = torch.empty_like(ud)
    run = lambda **kw: ops.matvec(blocks, Ed, ud, y, shape, mask=mask, **kw)  # noqa: E731
    run()
    assert rel(y, Au) < 1e-12
    fixed = m == 0
    assert np.array_equal(y.cpu().numpy()[fixed], u[fixed])
    run(fixed_zero=True)
    assert rel(y, Au * m) < 1e-12 and np.all(y.cpu().numpy()[fixed] == 0)
    run(mode=L.MODE_RESID, b=bd, fixed_zero=True)
    assert rel(y, (b - Au) * m) < 1e-12
    run(mode=L.MODE_RESID, b=bd)
    assert rel(y, b - Au) < 1e-12
    run(mode=L.MODE_JACOBI, b=bd, dinv=dd, omega=0.43)
    assert rel(y, u + 0.43 * dinv * (b - Au)) < 1e-12
    d = dev(d_old)
    run(mode=L.MODE_CHEB, b=bd, dinv=dd, d=d, cheb_a=0.7, cheb_c=0.2)
    dn = 0.7 * (dinv * (b - Au) + 0.2 * d_old)
    assert rel(d, dn) < 1e-12 and rel(y, u + dn) < 1e-12
    nt, _ = ops.matvec_tiles(shape, blocks=blocks)
    part = torch.zeros(nt, dtype=torch.float64, device="cuda")
    run(mode=L.MODE_DOT, partial=part)
    assert float(part.sum()) == pytest.approx(float(u @ Au), rel=1e-12)


@pytest.mark.parametrize("kind", ["none", "x0", "array"])
def test_diagonal_and_inverse_diagonal(kind):
    from voxelhip import ops
    shape = (7, 6, 9)
    rng = np.random.default_rng(2)
    E, _ = case(shape)
    prod = hx.NumpyProduct(shape, KE)
    d_ref = prod.diagonal(E)
    out = torch.empty(d_ref.size, dtype=torch.float64, device="cuda")
    ops.diagonal(KE, dev(E), out, shape)
    assert rel(out, d_ref) < 1e-14
    if kind != "none":
        m = _masks(shape, rng)[kind]
        inv_ref = mg_ref.Level(shape, prod, E, m, np.float64).inv_diag
        ops.diagonal(KE, dev(E), out, shape, mask=ops.Mask.from_values(m, shape), invert=True)
        assert rel(out, inv_ref) < 1e-14


def test_element_energy_and_sensitivity():
    from oracle import simpref as sr
    from voxelhip import ops
    shape = (9, 6, 10)
    rng = np.random.default_rng(3)
    rho = rng.uniform(0.01, 1, int(np.prod(shape)))
    u = rng.standard_normal(3 * hx.n_nodes(shape))
    ke = 2 * KE
    ce_ref = hx.element_energy(shape, ke, u)
    ce = torch.empty(rho.size, dtype=torch.float64, device="cuda")
    dc = torch.empty_like(ce)
    ops.element_energy(hx.Modal(ke).blocks, dev(u), ce, shape, rho=dev(rho), dc=dc, penal=3.0, dE=1 - 1e-9)
    assert rel(ce, ce_ref) < 1e-12
    assert rel(dc, sr.sensitivity(rho, ce_ref)) < 1e-12


@pytest.mark.parametrize("fine", [(4, 2, 2), (16, 8, 8), (34, 18, 66)])
def test_transfers_match_the_oracle(fine):
    import voxel_mg as mg
    rng = np.random.default_rng(4)
    coarse = tuple(s // 2 for s in fine)
    uc = rng.standard_normal(3 * hx.n_nodes(coarse))
    rf = rng.standard_normal(3 * hx.n_nodes(fine))
    assert rel(mg.prolong(None, dev(uc), coarse), mg_ref.prolong(uc, coarse)) < 1e-15
    assert rel(mg.restrict(None, dev(rf), fine), mg_ref.restrict(rf, fine)) < 1e-15
    E = rng.random(int(np.prod(fine)))
    assert rel(mg.coarsen_modulus(None, dev(E), fine), mg_ref.coarsen_modulus(E, fine)) < 1e-15


@pytest.mark.parametrize("kind", ["x0", "array"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("fine", [(6, 20, 340), (4, 10, 322), (10, 18, 16)])
def test_masked_transfers_over_row_blocks_and_z_segments(kind, dtype, fine):
    """Masked restriction and accumulating prolongation on grids whose rows span several
    blocks in y and several segments in z (340 and 322 fine elements along z)."""
    from voxelhip import _lib as L, ops
    rng = np.random.default_rng(11)
    coarse = tuple(s // 2 for s in fine)
    mf = _masks(fine, rng)[kind]
    mc = mg_ref.coarsen_mask(mf, fine)
    Mf, Mc = ops.Mask.from_values(mf, fine), ops.Mask.from_values(mc, coarse)
    want = L.MASK_X0 if kind == "x0" else L.MASK_ARRAY
    assert Mf.kind == want and Mc.kind == want
    rf = rng.standard_normal(mf.size)
    uc = rng.standard_normal(mc.size)
    xf = rng.standard_normal(mf.size)
    tol = 1e-15 if dtype == torch.float64 else 1e-6
    rc = ops.restrict(dev(rf, dtype), torch.empty(mc.size, dtype=dtype, device="cuda"), fine, fmask=Mf, cmask=Mc)
    assert rel(rc, mg_ref.restrict(rf * mf, fine) * mc) < tol
    out = ops.prolong(dev(uc, dtype), dev(xf, dtype), fine, mask=Mf, assign=False)
    assert rel(out, xf + mf * mg_ref.prolong(uc, coarse)) < tol
    out = ops.prolong(dev(uc, dtype), dev(xf, dtype), fine, mask=Mf, assign=True)
    assert rel(out, mf * mg_ref.prolong(uc, coarse)) < tol


def test_full_size_product_properties():
    """256 x 128 x 128 (12.8 M DOFs): the parity kernel against the independent dense-ke kernel,
    linearity, and rigid translations / rotations in the null space."""
    from voxelhip import ops
    shape = (256, 128, 128)
    print("case", shape, flush=True)
    g = torch.Generator(device="cuda").manual_seed(1)
    ne, nd = int(np.prod(shape)), 3 * hx.n_nodes(shape)
    E = 1e-9 + torch.rand(ne, generator=g, device="cuda", dtype=torch.float64) ** 3
    u = torch.randn(nd, generator=g, device="cuda", dtype=torch.float64)
    v = torch.randn(nd, generator=g, device="cuda", dtype=torch.float64)
    blocks = hx.Modal(KE).blocks
    y, yd, yv, yl = (torch.empty_like(u) for _ in range(4))
    ops.matvec(blocks, E, u, y, shape)
    ops.dense_product(KE, E, u, yd, shape)
    assert float((y - yd).norm() / yd.norm()) < 1e-12
    ops.matvec(blocks, E, v, yv, shape)
    ops.matvec(blocks, E, (2.5 * u - v).contiguous(), yl, shape)
    assert float((yl - (2.5 * y - yv)).norm() / yl.norm()) < 1e-12
    nx, ny, nz = shape
    X = torch.stack(torch.meshgrid(*[torch.arange(n + 1, device="cuda", dtype=torch.float64) for n in shape],
                                   indexing="ij"), -1).reshape(-1, 3)
    rot = torch.stack([-X[:, 1], X[:, 0], torch.zeros_like(X[:, 0])], 1).reshape(-1).contiguous()
    ops.matvec(blocks, E, rot, yl, shape)
    assert float(yl.abs().max()) < 1e-12 * float(rot.abs().max())


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("shape", [(1, 1, 1), (5, 3, 7), (16, 9, 12)])
def test_lanes24_control_matches_the_product(dtype, shape):
    """The element-per-24-lanes layout control (atomics, two elements per wave64) against the
    plane-sweep product and the oracle."""
    from voxelhip import ops
    print("case", dtype, shape, flush=True)
    g = torch.Generator(device="cuda").manual_seed(11)
    ne, nd = int(np.prod(shape)), 3 * hx.n_nodes(shape)
    E = torch.rand(ne, generator=g, device="cuda", dtype=torch.float64) + 0.05
    u = torch.randn(nd, generator=g, device="cuda", dtype=torch.float64)
    ref = hx.NumpyProduct(shape, KE).raw(u.cpu().numpy(), E.cpu().numpy())
    blocks = hx.Modal(KE).blocks
    y = torch.full((nd,), 7.0, dtype=dtype, device="cuda")          # not zero on entry
    ops.lanes24_product(blocks, E.to(dtype), u.to(dtype), y, shape)
    yp = ops.matvec(blocks, E.to(dtype), u.to(dtype), torch.empty_like(y), shape)
    tol = 1e-12 if dtype == torch.float64 else 2e-6
    assert np.linalg.norm(y.double().cpu().numpy() - ref) / np.linalg.norm(ref) < tol
    assert float((y - yp).double().norm() / yp.double().norm()) < tol


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_dense_control_is_cached_and_graph_capturable(dtype):
    """The node-owned dense control keeps its ke on the device per distinct matrix: after one eager
    call it replays inside a HIP graph (every replay equal to the eager product); a second ke is its
    own entry; a new ke inside a capture is refused."""
    from voxelhip import ops
    shape = (12, 6, 10)
    g = torch.Generator(device="cuda").manual_seed(7)
    ne, nd = int(np.prod(shape)), 3 * hx.n_nodes(shape)
    E = (torch.rand(ne, generator=g, device="cuda", dtype=torch.float64) + 0.1).to(dtype)
    u = torch.randn(nd, generator=g, device="cuda", dtype=torch.float64).to(dtype)
    ref = ops.dense_product(KE, E, u, torch.empty_like(u), shape).clone()
    ref2 = ops.dense_product(2 * KE, E, u, torch.empty_like(u), shape).clone()
    assert torch.allclose(ref2, 2 * ref, rtol=1e-12 if dtype == torch.float64 else 1e-5, atol=0)
    y = torch.empty_like(u)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        ops.dense_product(KE, E, u, y, shape)
    for _ in range(3):
        y.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, ref)
    del graph
    with warnings.catch_warnings():
        warnings.filterwarnings("ignore", message="The CUDA Graph is empty")   # the refused capture records nothing
        with pytest.raises(Exception):
            g2 = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g2):
                ops.dense_product(3.5 * KE, E, u, y, shape)
    torch.cuda.synchronize()


def test_product_and_fused_passes_are_reproducible_under_repetition():
    """Regression for LDS ordering: many launches of every epilogue must give bitwise-identical
    output (a missing barrier between the staged-plane reads and the next plane's writes shows
    up as run-to-run differences)."""
    from voxelhip import _lib as L, ops
    shape = (64, 48, 96)
    E, u = case(shape, 8)
    m = np.ones((shape[0] + 1, (shape[1] + 1) * (shape[2] + 1) * 3)); m[0] = 0
    mask = ops.Mask.from_values(m.ravel(), shape)
    blocks = hx.Modal(KE).blocks
    ref = hx.NumpyProduct(shape, KE).raw(u, E)
    for dt in (torch.float64, torch.float32):
        Ed, ud = dev(E, dt), dev(u, dt)
        first = torch.empty_like(ud)
        ops.matvec(blocks, Ed, ud, first, shape)
        assert rel(first, ref) < (1e-12 if dt == torch.float64 else 1e-5)
        y = torch.empty_like(ud)
        for _ in range(30):
            ops.matvec(blocks, Ed, ud, y, shape)
            assert torch.equal(y, first)
        ops.matvec(blocks, Ed, ud, first, shape, mask=mask, mode=L.MODE_RESID, b=ud)
        for _ in range(10):
            ops.matvec(blocks, Ed, ud, y, shape, mask=mask, mode=L.MODE_RESID, b=ud)
            assert torch.equal(y, first)


@pytest.mark.parametrize("kind", ["x0", "array"])
def test_three_term_chebyshev_and_diagonal_from_moduli(kind):
    from voxelhip import _lib as L, ops
    shape = (18, 17, 34)
    rng = np.random.default_rng(12)
    E, u = case(shape, 6)
    m = _masks(shape, rng)[kind]
    prod = hx.NumpyProduct(shape, KE)
    lvl = mg_ref.Level(shape, prod, E, m, np.float64)
    Au = lvl.apply(u)
    b = rng.standard_normal(u.size)
    prev = rng.standard_normal(u.size)
    dinv = lvl.inv_diag
    kd = np.diag(KE)[:3]
    mask = ops.Mask.from_values(m, shape)
    blocks = hx.Modal(KE).blocks
    Ed, ud, bd = dev(E), dev(u), dev(b)
    for dk in ({"dinv": dev(dinv)}, {"kdiag": kd}):
        y = dev(prev)
        ops.matvec(blocks, Ed, ud, y, shape, mask=mask, mode=L.MODE_CHEB3, b=bd, cheb_a=0.7, cheb_c=0.2, **dk)
        assert rel(y, u + 0.7 * (dinv * (b - Au) + 0.2 * (u - prev))) < 1e-12
        y = dev(prev)
        ops.matvec(blocks, Ed, ud, y, shape, mask=mask, mode=L.MODE_CHEB3, b=bd, cheb_a=0.7, cheb_c=0.2,
                   prev_zero=True, **dk)
        assert rel(y, u + 0.7 * (dinv * (b - Au) + 0.2 * u)) < 1e-12
        y = torch.full_like(ud, float("nan"))
        ops.matvec(blocks, Ed, ud, y, shape, mask=mask, mode=L.MODE_CHEB3, b=bd, cheb_a=0.7, cheb_c=0.0, **dk)
        assert rel(y, u + 0.7 * dinv * (b - Au)) < 1e-12
        ops.matvec(blocks, Ed, ud, y, shape, mask=mask, mode=L.MODE_JACOBI, b=bd, omega=0.43, **dk)
        assert rel(y, u + 0.43 * dinv * (b - Au)) < 1e-12


def test_fp32_product_takes_the_cube_path_and_matches_fp64_blocks():
    """The FP32 product keeps the voxel-cube fast path (blocks from the FP64 element matrix)."""
    import voxel_mg as mg
    shape = (20, 18, 40)
    E, u = case(shape, 9)
    p32 = mg.HipProduct(shape, KE, np.float32)
    from voxelhip import ops
    assert ops.matvec_tiles(shape, dtype=torch.float32, blocks=p32.blocks) == \
        ops.matvec_tiles(shape, dtype=torch.float32, blocks=None)
    y = p32.raw(dev(u, torch.float32), dev(E, torch.float32))
    assert rel(y, hx.NumpyProduct(shape, KE).raw(u, E)) < 1e-5


@pytest.mark.parametrize("kind", ["none", "x0", "array"])
@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4])
def test_resid_pass_reads_an_fp32_direction(kind, tile):
    """FP64 RESID pass with the AXPY load reading an FP32 w (the mixed-precision Krylov loop's
    search direction): bitwise the pass fed the FP64 image of the same values, and the oracle's
    x + alpha p and b - A x; an FP32 w anywhere else is refused."""
    from voxelhip import _lib as L, ops
    shape = (20, 18, 36)
    rng = np.random.default_rng(31)
    E, x_old = case(shape, 12)
    m = np.ones(3 * hx.n_nodes(shape)) if kind == "none" else _masks(shape, rng)[kind]
    lvl = mg_ref.Level(shape, hx.NumpyProduct(shape, KE), E, m, np.float64)
    p32 = rng.standard_normal(x_old.size).astype(np.float32)
    b = rng.standard_normal(x_old.size)
    alpha = -0.8125
    scal = torch.zeros(16, dtype=torch.float64, device="cuda")
    scal[L.S_ALPHA] = alpha
    mask = ops.Mask.from_values(m, shape)
    blocks = hx.Modal(KE).blocks
    Ed, xd, bd = dev(E), dev(x_old), dev(b)
    nt, _ = ops.matvec_tiles(shape, blocks=blocks, tile=tile)
    outs = []
    print("case", kind, tile, flush=True)
    for w in (dev(p32, torch.float32), dev(p32.astype(np.float64))):
        part = torch.zeros(nt, dtype=torch.float64, device="cuda")
        r, xn = torch.empty_like(xd), torch.empty_like(xd)
        r32 = torch.empty(x_old.size, dtype=torch.float32, device="cuda")
        ops.matvec(blocks, Ed, xd, r, shape, mask=mask, mode=L.MODE_RESID, b=bd, partial=part, w=w, sdev=scal,
                   sidx=L.S_ALPHA, u_out=xn, y32=r32, tile=tile)
        outs.append((r, xn, r32, part))
    for a, c in zip(outs[0], outs[1]):
        assert torch.equal(a, c)
    r, xn, r32, part = outs[0]
    x_ref = x_old + alpha * p32.astype(np.float64)
    r_ref = b - lvl.apply(x_ref)
    assert rel(xn, x_ref) < 1e-15 and rel(r, r_ref) < 1e-12
    assert float(part.sum()) == pytest.approx(float(r_ref @ r_ref), rel=1e-12)
    with pytest.raises(RuntimeError, match="FP32 w"):   # an FP32 w in the FP64 DOT pass
        ops.matvec(blocks, Ed, xd, None, shape, mask=mask, mode=L.MODE_DOT, partial=part,
                   w=dev(p32, torch.float32), sdev=scal, sidx=L.S_BETA, u_out=torch.empty_like(xd))


def test_fused_krylov_passes():
    """DOT pass with p = z + beta p_old formed on load, RESID pass with x = x_old + alpha p formed
    on load (+ FP32 copy and |r|^2 partials), and the convert + Polak-Ribiere-dots kernel."""
    from voxelhip import _lib as L, ops
    shape = (20, 18, 36)
    rng = np.random.default_rng(21)
    E, z = case(shape, 10)
    m = _masks(shape, rng)["x0"]
    lvl = mg_ref.Level(shape, hx.NumpyProduct(shape, KE), E, m, np.float64)
    p_old, x_old, b = (rng.standard_normal(z.size) for _ in range(3))
    beta, alpha = 0.37, -1.25
    scal = torch.zeros(16, dtype=torch.float64, device="cuda")
    scal[L.S_BETA], scal[L.S_ALPHA] = beta, alpha
    mask = ops.Mask.from_values(m, shape)
    blocks = hx.Modal(KE).blocks
    Ed = dev(E)
    nt, _ = ops.matvec_tiles(shape, blocks=blocks)
    part = torch.zeros(nt, dtype=torch.float64, device="cuda")
    q, pn = torch.empty_like(Ed.new_empty(z.size)), torch.empty(z.size, dtype=torch.float64, device="cuda")
    ops.matvec(blocks, Ed, dev(z), q, shape, mask=mask, mode=L.MODE_DOT, partial=part, w=dev(p_old), sdev=scal,
               sidx=L.S_BETA, u_out=pn)
    p_ref = z + beta * p_old
    Ap = lvl.apply(p_ref)
    assert rel(pn, p_ref) < 1e-15 and rel(q, Ap) < 1e-12
    assert float(part.sum()) == pytest.approx(float(p_ref @ Ap), rel=1e-12)
    r, xn = torch.empty_like(pn), torch.empty_like(pn)
    r32 = torch.empty(z.size, dtype=torch.float32, device="cuda")
    ops.matvec(blocks, Ed, dev(x_old), r, shape, mask=mask, mode=L.MODE_RESID, b=dev(b), partial=part,
               w=pn, sdev=scal, sidx=L.S_ALPHA, u_out=xn, y32=r32)
    x_ref = x_old + alpha * p_ref
    r_ref = b - lvl.apply(x_ref)
    assert rel(xn, x_ref) < 1e-15 and rel(r, r_ref) < 1e-12
    assert float(part.sum()) == pytest.approx(float(r_ref @ r_ref), rel=1e-12)
    assert rel(r32, r_ref) < 1e-7
    zin = rng.standard_normal(z.size).astype(np.float32)
    zold = rng.standard_normal(z.size)
    znew = torch.empty_like(pn)
    nparts = ops.reduce_parts(z.size)
    p2 = torch.zeros(2 * nparts, dtype=torch.float64, device="cuda")
    ops.cvt_dotdiff(dev(zin, torch.float32), r, dev(zold), znew, p2)
    zd = zin.astype(np.float64)
    assert torch.equal(znew, dev(zd))
    assert float(p2[:nparts].sum()) == pytest.approx(float(r_ref @ zd), rel=1e-12)
    assert float(p2[nparts:].sum()) == pytest.approx(float(r_ref @ (zd - zold)), rel=1e-12)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_chebyshev_step_with_fused_polak_ribiere_dots(dtype):
    """CHEB3D: the three-term step with the previous iterate read from `prev` writes the same y
    as CHEB3 in place, and its per-tile partials sum to b . y and b . (y - zold)."""
    from voxelhip import _lib as L, ops
    shape = (18, 17, 34)
    rng = np.random.default_rng(31)
    E, u = case(shape, 8)
    m = _masks(shape, rng)["x0"]
    mask = ops.Mask.from_values(m, shape)
    blocks = hx.Modal(KE).blocks
    kd = np.diag(KE)[:3]
    b, prev, zold = (rng.standard_normal(u.size) for _ in range(3))
    Ed, ud, bd, pd, zd = dev(E, dtype), dev(u, dtype), dev(b, dtype), dev(prev, dtype), dev(zold, dtype)
    ref = pd.clone()
    ops.matvec(blocks, Ed, ud, ref, shape, mask=mask, mode=L.MODE_CHEB3, b=bd, cheb_a=0.7, cheb_c=0.2, kdiag=kd)
    nt = ops.matvec_tiles(shape, dtype=dtype, blocks=blocks)[0]
    partial = torch.full((2 * nt,), float("nan"), dtype=torch.float64, device="cuda")
    y = torch.full_like(ud, float("nan"))
    ops.matvec(blocks, Ed, ud, y, shape, mask=mask, mode=L.MODE_CHEB3D, b=bd, cheb_a=0.7, cheb_c=0.2, kdiag=kd,
               prev=pd, zold=zd, partial=partial)
    assert torch.equal(y, ref)
    yh, bh, zh = (t.double().cpu().numpy() for t in (y, bd, zd))
    tol = 1e-12 if dtype == torch.float64 else 1e-9
    assert float(partial[:nt].sum()) == pytest.approx(float(bh @ yh), rel=tol)
    assert float(partial[nt:].sum()) == pytest.approx(float(bh @ (yh - zh)), rel=tol)
    # prev given for a plain CHEB3 step: y is then only written
    y2 = torch.full_like(ud, float("nan"))
    ops.matvec(blocks, Ed, ud, y2, shape, mask=mask, mode=L.MODE_CHEB3, b=bd, cheb_a=0.7, cheb_c=0.2, kdiag=kd,
               prev=pd)
    assert torch.equal(y2, ref)


def test_dot_pass_reads_an_fp32_input():
    """The FP64 DOT pass with the AXPY load reads an FP32 z exactly as its FP64 conversion."""
    from voxelhip import _lib as L, ops
    shape = (20, 12, 30)
    rng = np.random.default_rng(32)
    E, _ = case(shape, 9)
    m = _masks(shape, rng)["x0"]
    mask = ops.Mask.from_values(m, shape)
    blocks = hx.Modal(KE).blocks
    z32 = torch.from_numpy(rng.standard_normal(3 * hx.n_nodes(shape))).to("cuda", torch.float32)
    pold = dev(rng.standard_normal(z32.numel()))
    sdev = torch.zeros(16, dtype=torch.float64, device="cuda")
    sdev[L.S_BETA] = 0.37
    nt = ops.matvec_tiles(shape, blocks=blocks)[0]
    out = []
    for z in (z32, z32.double()):
        partial = torch.zeros(nt, dtype=torch.float64, device="cuda")
        p = torch.empty_like(pold)
        ops.matvec(blocks, dev(E), z, None, shape, mask=mask, mode=L.MODE_DOT, partial=partial, w=pold, sdev=sdev,
                   sidx=L.S_BETA, u_out=p)
        out.append((p, partial))
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])
    ph = out[1][0].cpu().numpy()
    assert np.allclose(ph, z32.double().cpu().numpy() + 0.37 * pold.cpu().numpy(), rtol=1e-15, atol=1e-15)


def test_resid_pass_may_keep_only_its_fp32_copy():
    from voxelhip import _lib as L, ops
    shape = (20, 12, 30)
    rng = np.random.default_rng(33)
    E, u = case(shape, 10)
    mask = ops.Mask.from_values(_masks(shape, rng)["x0"], shape)
    blocks = hx.Modal(KE).blocks
    b = dev(rng.standard_normal(u.size))
    nt = ops.matvec_tiles(shape, blocks=blocks)[0]
    outs = []
    for keep in (True, False):
        y = torch.empty_like(b) if keep else None
        y32 = torch.full((u.size,), float("nan"), dtype=torch.float32, device="cuda")
        partial = torch.zeros(nt, dtype=torch.float64, device="cuda")
        ops.matvec(blocks, dev(E), dev(u), y, shape, mask=mask, mode=L.MODE_RESID, b=b, partial=partial, y32=y32)
        outs.append((y, y32, partial))
    assert torch.equal(outs[0][1], outs[1][1]) and torch.equal(outs[0][2], outs[1][2])
    assert torch.equal(outs[0][0].float(), outs[1][1])
