"""Pins the element algebra and the product of the oracle (oracle/hexref.py) against the
reference's fixtures, the reference's own C++ oracle (oracle/_ref) and closed forms."""
import numpy as np
import pytest

from oracle import cref
from oracle import hexref as hx

SHAPE = (3, 2, 4)


def _case(seed=7):
    """tests/test_kernels_cpu.py:14-20 of the reference."""
    rng = np.random.default_rng(seed)
    ke = hx.quadrature([1.0, 1.0, 1.0], 0.3)
    E = 1e-6 + (1 - 1e-6) * rng.random(int(np.prod(SHAPE))) ** 3
    u = rng.standard_normal(3 * hx.n_nodes(SHAPE))
    return ke, E, u


def test_unit_cube_matches_the_reference_fixtures(golden):
    q = hx.quadrature([1, 1, 1], 0.3)
    ku = np.load(golden / "ke_unit_3d.npy")
    # the donor's KE_UNIT_3D convention is twice the unit-cube integral (a cube of side 2)
    assert np.abs(ku - 2 * q).max() < 1e-15
    smoke = np.load(golden / "smoke-q4.npz")
    assert np.abs(smoke["ke"] * 8 - ku).max() < 1e-15           # fixture ke = unit / nx
    opt = np.load(golden / "optimized-q64.npz")
    assert np.abs(opt["ke"] * 128 - ku).max() < 1e-14


@pytest.mark.skipif(not cref.available(), reason="oracle/_ref not built (needs /root/reference)")
def test_quadrature_matches_the_reference_cpp_element():
    R = cref.Reference()
    for h in (1.0, 0.25, 3.0):
        assert np.abs(R.element(h, 0.3) - hx.quadrature([h, h, h], 0.3)).max() < 1e-15 * max(1, h)


def test_quadrature_closed_forms():
    lam, mu = 0.3 / (1.3 * 0.4), 1 / 2.6
    for lengths in ([1, 1, 1], [1, 2, 0.5]):
        ke = hx.quadrature(lengths, 0.3)
        assert np.abs(ke - ke.T).max() < 1e-15
        ev = np.linalg.eigvalsh(ke)
        assert np.all(np.abs(ev[:6]) < 1e-14) and ev[6] > 1e-3      # exactly 6 rigid modes
        X = hx.XYZ * np.asarray(lengths)
        # uniform strain eps_xx = 1: energy u^T K u = (lam + 2 mu) * volume
        u = np.zeros(24); u[0::3] = X[:, 0]
        assert u @ ke @ u == pytest.approx((lam + 2 * mu) * np.prod(lengths), rel=1e-13)
        # pure shear gamma_xy = 1: energy = mu * volume
        u = np.zeros(24); u[0::3] = X[:, 1]
        assert u @ ke @ u == pytest.approx(mu * np.prod(lengths), rel=1e-13)
        # rotation about z is a null mode
        u = np.zeros(24); u[0::3], u[1::3] = -X[:, 1], X[:, 0]
        assert np.abs(ke @ u).max() < 1e-14


def test_modal_factorisation_closed_forms():
    assert np.abs(hx.T24 @ hx.T24.T - 8 * np.eye(24)).max() == 0
    for lengths in ([1, 1, 1], [1, 2, 0.5]):
        ke = hx.quadrature(lengths, 0.3)
        m = hx.Modal(ke)
        assert m.discarded_relative < 2e-14
        assert np.abs(m.reconstructed - ke).max() < 1e-14
        v = np.random.default_rng(0).standard_normal((7, 24))
        assert np.abs(m.apply(v) - v @ ke.T).max() < 1e-14


def test_modal_rejects_a_non_brick_element():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((24, 24))
    with pytest.raises(ValueError):
        hx.Modal(a @ a.T)


@pytest.mark.parametrize("path", ["modal8", "dense8"])
def test_fp64_lane_emulation_matches_assembled_product(path):
    ke, E, u = _case()
    ref = hx.assembled_reference(SHAPE, ke, E, u)
    y = hx.emulate(path, SHAPE, ke, E, u, np.float64)
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-13


@pytest.mark.parametrize("path", ["modal8", "dense8"])
def test_fp32_lane_emulation_accuracy(path):
    ke, E, u = _case()
    ref = hx.assembled_reference(SHAPE, ke, E, u)
    y = hx.emulate(path, SHAPE, ke, E, u, np.float32)
    assert y.dtype == np.float32
    assert np.linalg.norm(y.astype(np.float64) - ref) / np.linalg.norm(ref) < 1e-5


def test_dropped_butterfly_is_detected(monkeypatch):
    """Negative control (tests/test_kernels_cpu.py:53-63): a planted mistake in the oracle's
    lane transform must break agreement with the brute-force product."""
    ke, E, u = _case()
    ref = hx.assembled_reference(SHAPE, ke, E, u)
    monkeypatch.setattr(hx, "lane_transform", lambda v: v)
    y = hx.emulate("modal8", SHAPE, ke, E, u)
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) > 1e-2


def test_assembled_reference_matches_dense_assembly():
    ke, E, u = _case()
    A = hx.assemble_dense(SHAPE, ke, E, np.ones(3 * hx.n_nodes(SHAPE)))
    assert np.linalg.norm(hx.assembled_reference(SHAPE, ke, E, u) - A @ u) < 1e-13 * np.linalg.norm(A @ u)


def test_numpy_product_and_diagonal():
    ke, E, u = _case()
    p = hx.NumpyProduct(SHAPE, ke)
    ref = hx.assembled_reference(SHAPE, ke, E, u)
    assert np.linalg.norm(p.raw(u, E) - ref) / np.linalg.norm(ref) < 1e-14
    A = hx.assemble_dense(SHAPE, ke, E, np.ones(3 * hx.n_nodes(SHAPE)))
    assert np.abs(p.diagonal(E) - np.diag(A)).max() < 1e-15


def test_rigid_body_modes_are_in_the_null_space():
    ke, E, _ = _case()
    nx, ny, nz = SHAPE
    X = np.stack(np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1),
                             indexing="ij"), -1).reshape(-1, 3).astype(float)
    p = hx.NumpyProduct(SHAPE, ke)
    for u in (np.tile([1.0, 0, 0], len(X)),
              np.stack([-X[:, 1], X[:, 0], 0 * X[:, 0]], 1).ravel(),
              np.stack([0 * X[:, 0], -X[:, 2], X[:, 1]], 1).ravel()):
        assert np.abs(p.raw(u, E)).max() < 1e-13


def test_layout_round_trip():
    shape = (3, 2, 4)
    a = np.arange(3 * hx.n_nodes(shape), dtype=float)
    assert np.array_equal(hx.to_natural(hx.to_author(a, shape, True), shape, True), a)
    b = np.arange(24.0)
    assert np.array_equal(hx.to_author(hx.to_natural(b, shape), shape), b)


@pytest.mark.skipif(not cref.available(), reason="oracle/_ref not built (needs /root/reference)")
def test_natural_product_matches_the_reference_cpp_oracle(golden):
    d = np.load(golden / "smoke-q4.npz")
    shape = tuple(int(s) for s in d["shape"])
    u = np.random.default_rng(3).standard_normal(d["mask"].size)
    y, ce, dg = cref.Reference().product(shape, d["rho"], u, d["ke"], d["mask"], 1e-6)
    y2, ce2, dg2 = hx.natural_product(shape, d["rho"], u, d["ke"], d["mask"], 1e-6, 1.0, 3.0)
    assert np.linalg.norm(y2 - y) / np.linalg.norm(y) < 1e-14
    assert np.linalg.norm(ce2 - ce) / np.linalg.norm(ce) < 1e-14
    assert np.abs(dg2 - dg).max() < 1e-16
