"""vh_matvec_dot_noalloc (the DOT pass with its partial sums in the library's own device array,
no workspace, no allocation) against the existing entry point vh_matvec_fused (DOT pass with a
caller-allocated partial array) + vh_finish, through the C-ABI: y and u.Au bitwise equal (same
kernels, same plan, same summation order), on two calls in a row with the second on a smaller
grid, in FP64 and FP32 and with each mask kind; grids past the array's size are rejected."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import hexref as hx

pytestmark = pytest.mark.gpu

KE = 2 * hx.quadrature([1, 1, 1], 0.3)
MAX_TILES = 65536


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _inputs(shape, dtype, seed):
    from voxelhip import ops
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    E = torch.from_numpy(rng.uniform(0.01, 1.0, nx * ny * nz)).to("cuda", dtype)
    u = torch.from_numpy(rng.standard_normal(3 * ops.n_nodes(shape))).to("cuda", dtype)
    return E, u


def _reference(lib, L, dtype_code, blocks, E, u, shape, mask_kind, mask):
    """DOT pass of vh_matvec_fused with a caller-allocated partial array, then vh_finish."""
    nt, nw = ctypes.c_int(), ctypes.c_int()
    L.check(lib.vh_matvec_tiles(dtype_code, dtype_code, blocks.ctypes.data, *shape, 0, 0, ctypes.byref(nt),
                                ctypes.byref(nw)), "tiles")
    partial = torch.zeros(nt.value, dtype=torch.float64, device="cuda")
    y = torch.empty_like(u)
    ep = L.Epilogue()
    ep.mode = L.MODE_DOT
    ep.partial = partial.data_ptr()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L.check(lib.vh_matvec_fused(dtype_code, dtype_code, blocks.ctypes.data, _ptr(E), _ptr(u), _ptr(y), *shape,
                                mask_kind, _ptr(mask), ctypes.byref(ep), st), "vh_matvec_fused")
    pq = torch.zeros(16, dtype=torch.float64, device="cuda")
    L.check(lib.vh_finish(_ptr(partial), nt.value, 1, _ptr(pq), 0, L.OP_NONE, st), "vh_finish")
    return y, pq[0].clone(), nt.value


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("mask_kind", ["none", "x0", "array"])
def test_noalloc_dot_pass_equals_the_workspace_entry_point(dtype, mask_kind):
    import voxel_kernels as vk
    from voxelhip import _lib as L
    from voxelhip import ops
    lib = L.load()
    code = L.F64 if dtype == torch.float64 else L.F32
    blocks = np.ascontiguousarray(vk.Modal(KE).blocks, dtype=np.float64).reshape(72)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for call, shape in enumerate(((96, 48, 40), (24, 16, 8))):       # the second call is smaller
        print("case", dtype, mask_kind, shape, flush=True)
        E, u = _inputs(shape, dtype, seed=call)
        n = u.numel()
        if mask_kind == "none":
            kind, mask = L.MASK_NONE, None
        elif mask_kind == "x0":
            kind, mask = L.MASK_X0, None
        else:
            m = np.random.default_rng(5 + call).random(n) > 0.1
            kind, mask = L.MASK_ARRAY, torch.from_numpy(m.astype(np.uint8)).cuda()
        y_ref, pq_ref, tiles = _reference(lib, L, code, blocks, E, u, shape, kind, mask)
        assert tiles <= MAX_TILES
        y = torch.full_like(u, float("nan"))
        pq = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
        L.check(lib.vh_matvec_dot_noalloc(code, blocks.ctypes.data, _ptr(E), _ptr(u), _ptr(y), *shape, kind,
                                          _ptr(mask), _ptr(pq), st), "vh_matvec_dot_noalloc")
        torch.cuda.synchronize()
        assert torch.equal(y, y_ref)
        assert float(pq) == float(pq_ref)
        # and the dot is u . A u
        assert float(pq) == pytest.approx(float((u.double() * y_ref.double()).sum()), rel=1e-12 if code == L.F64 else 1e-5)
        # the dot alone (y = NULL)
        pq2 = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
        L.check(lib.vh_matvec_dot_noalloc(code, blocks.ctypes.data, _ptr(E), _ptr(u), None, *shape, kind,
                                          _ptr(mask), _ptr(pq2), st), "vh_matvec_dot_noalloc")
        torch.cuda.synchronize()
        assert float(pq2) == float(pq_ref)


def test_noalloc_rejects_grids_beyond_its_array():
    import voxel_kernels as vk
    from voxelhip import _lib as L
    lib = L.load()
    blocks = np.ascontiguousarray(vk.Modal(KE).blocks, dtype=np.float64).reshape(72)
    shape = (1, 40000, 2048)
    nt, nw = ctypes.c_int(), ctypes.c_int()
    L.check(lib.vh_matvec_tiles(L.F64, L.F64, blocks.ctypes.data, *shape, 0, 0, ctypes.byref(nt), ctypes.byref(nw)),
            "tiles")
    assert nt.value > MAX_TILES          # the call below is rejected before any launch
    E = torch.zeros(8, dtype=torch.float64, device="cuda")
    pq = torch.zeros(1, dtype=torch.float64, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    r = lib.vh_matvec_dot_noalloc(L.F64, blocks.ctypes.data, _ptr(E), _ptr(E), None, *shape, L.MASK_NONE, None,
                                  _ptr(pq), st)
    assert r == -2 and b"exceed" in lib.vh_last_error()
