"""CPU checks of the C-ABI library: it loads, exports every symbol include/voxelhip.h declares,
and its host-only entry point (parity-block factorisation) agrees with the oracle."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "voxelhip.h"


def _declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(vh_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from voxelhip import _lib
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES)
    assert lib.vh_abi_version() == _lib.ABI_VERSION == 4


def test_modal_blocks_match_the_oracle():
    from oracle import hexref as hx
    from voxelhip import ops
    for lengths in ([1, 1, 1], [1, 2, 0.5]):
        ke = hx.quadrature(lengths, 0.3)
        b, disc = ops.modal_blocks(ke)
        ref = hx.Modal(ke)
        assert np.abs(b - ref.blocks).max() < 1e-15
        assert disc == pytest.approx(ref.discarded_relative, abs=1e-16)


def test_modal_blocks_are_safe_from_concurrent_threads():
    """ctypes releases the GIL around the call: eight threads factorising different bricks at once
    must each get exactly their single-threaded result (no shared scratch in the library)."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import hexref as hx
    from voxelhip import ops
    kes = [hx.quadrature([1.0, 1.0 + 0.25 * i, 2.0 - 0.125 * i], 0.2 + 0.02 * i) for i in range(8)]
    want = [ops.modal_blocks(ke)[0] for ke in kes]

    def run(i):
        return all(np.array_equal(ops.modal_blocks(kes[i])[0], want[i]) for _ in range(300))

    with ThreadPoolExecutor(8) as ex:
        assert all(ex.map(run, range(8)))


def test_modal_blocks_reject_a_non_brick():
    from voxelhip import ops
    a = np.random.default_rng(1).standard_normal((24, 24))
    with pytest.raises(ValueError, match="parity-block"):
        ops.modal_blocks(a @ a.T)


def test_product_quadrature_matches_the_oracle(golden):
    from oracle import hexref as hx
    import voxel_kernels as vk
    for lengths in ([1, 1, 1], [1, 2, 0.5], [0.25, 0.25, 0.25]):
        assert np.abs(vk.quadrature(lengths, 0.3) - hx.quadrature(lengths, 0.3)).max() < 1e-15
    assert np.abs(2 * vk.quadrature([1, 1, 1], 0.3) - np.load(golden / "ke_unit_3d.npy")).max() < 1e-15


def test_make_source_returns_the_hip_kernel_source(tmp_path):
    """'double' returns PROTOTYPE_SOURCE, 'float' a different text that instantiates the FP32
    kernels only; both are self-contained (they pass hipcc's front end for gfx950 alone)."""
    import shutil
    import subprocess
    import voxel_kernels as vk
    src = vk.make_source("double")
    assert src is vk.PROTOTYPE_SOURCE and "parity_xmarch" in src
    f32 = vk.make_source("float")
    assert f32 != src and "parity_xmarch" in f32
    assert "#define VH_MATVEC_F32 1" in f32 and "#define VH_MATVEC_F64 1" not in f32
    assert "#define VH_MATVEC_F64 1" in src and "#define VH_MATVEC_F32 1" not in src
    assert '#include "' not in src and '#include "' not in f32
    with pytest.raises(ValueError):
        vk.make_source("half")
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("no hipcc")
    for name, text in (("f64.hip", src), ("f32.hip", f32)):
        (tmp_path / name).write_text(text)
        r = subprocess.run([hipcc, "-std=c++17", "--offload-arch=gfx950", "--cuda-device-only", "-fsyntax-only",
                            str(tmp_path / name)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]


def test_device_entry_points_fail_loudly_without_a_gpu():
    import torch
    import voxel_kernels as vk
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(RuntimeError):
        vk.emulate("modal8", (2, 2, 2), vk.quadrature([1, 1, 1], 0.3), np.ones(8), np.zeros(81))


def test_first_chebyshev_step_must_start_from_zero():
    """The CHEB3 pass with the load-time product is the first step from a zero iterate (it does not
    read the previous one): the entry points reject it unless VH_PREV_ZERO is set or cheb_c = 0.
    (Argument checks only: the rejected calls return before touching a device.)"""
    from voxelhip import _lib as L
    lib = L.load()
    fake = ctypes.c_void_p(4096)
    blocks = (ctypes.c_double * 72)()
    ep = L.Epilogue()
    ep.mode = L.MODE_CHEB3
    ep.b, ep.w, ep.u_out = 8192, 12288, 16384
    ep.cheb_a, ep.cheb_c, ep.load_scale = 0.5, 0.3, 1.0
    ep.kdiag[0] = ep.kdiag[1] = ep.kdiag[2] = 1.0
    args = (L.F64, L.F64, ctypes.addressof(blocks), fake, fake, ctypes.c_void_p(20480), 4, 4, 4, L.MASK_X0, None, ctypes.byref(ep), None)
    assert lib.vh_matvec_fused(*args) == -2
    assert "first step from a zero iterate" in lib.vh_last_error().decode()
    item = L.ProductItem()
    item.blocks72, item.E, item.u, item.y = ctypes.addressof(blocks), 4096, 4096, 20480
    item.nx = item.ny = item.nz = 4
    item.ep = ep
    h = ctypes.c_void_p()
    assert lib.vh_product_batch_create(L.F64, L.MASK_X0, 0, 1, ctypes.byref(item), ctypes.byref(h), None) == -2
    assert "starts from zero" in lib.vh_last_error().decode()


def test_host_index_helpers_match_the_oracle():
    from oracle import hexref as hx
    import voxel_kernels as vk
    import voxel_mg as mg
    for shape in ((3, 2, 4), (1, 1, 1), (5, 3, 2)):
        assert np.array_equal(vk.element_nodes(shape), hx.element_nodes(shape))
        assert np.array_equal(mg.edof_author(shape), hx.edof_author(shape))
        assert mg.hierarchy_shapes((128, 64, 64)) == [(128, 64, 64), (64, 32, 32), (32, 16, 16), (16, 8, 8)]


def test_assemble_dense_matches_the_oracle():
    from oracle import hexref as hx
    import voxel_mg as mg
    rng = np.random.default_rng(0)
    shape = (3, 2, 2)
    ke = hx.quadrature([1, 1, 1], 0.3)
    E = rng.random(12)
    mask = (rng.random(3 * hx.n_nodes(shape)) > 0.2).astype(float)
    assert np.abs(mg.assemble_dense(shape, ke, E, mask) - hx.assemble_dense(shape, ke, E, mask)).max() < 1e-15


def test_product_continuation_schedule_matches_the_donor_module(golden):
    """voxelhip.simp.schedule_at against the donor's simp_r2.schedule_at (tests/golden/schedule_golden.json)."""
    import json
    from voxelhip import simp
    g = json.loads((golden / "schedule_golden.json").read_text())
    for it, row in enumerate(g["default"], start=1):
        assert list(simp.schedule_at(it)) == row
    table = [tuple(r) for r in g["custom_table"]]
    for it, row in enumerate(g["custom"], start=1):
        assert list(simp.schedule_at(it, table)) == row
