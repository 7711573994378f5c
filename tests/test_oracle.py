"""Pins of the CPU oracle (oracle/le2d_ref.c) before it is trusted as the checker.

1. Golden vectors written by the compiled reference (tests/golden): every
   oracle function on the step path must reproduce them bit for bit.
2. Closed forms and invariants that a plausible slip (a dropped eps, a swapped
   invariant, a wrong sign, a transposed stencil) would break.
"""
import math
import os
import tempfile

import numpy as np
import pytest

import golden_data as gd
from oracle import EDGE_CLAMP, EDGE_RING, Oracle

OD = Oracle(np.float64)
OF = Oracle(np.float32)


# --------------------------------------------------------------------------
# 1. golden vectors


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_tiny_golden(dtype):
    o = OD if dtype == np.float64 else OF
    cases, data = gd.tiny(dtype)
    for c in cases:
        k, dt, h, mat = c["key"], c["dt"], c["h"], c["mat"]
        g0 = data[k + "_init"]
        g = g0.copy()
        o.sweep_x(g, dt, h[0], mat)
        assert np.array_equal(g, data[k + "_x"]), k + " x"
        g = g0.copy()
        o.sweep_y(g, dt, h[1], mat)
        assert np.array_equal(g, data[k + "_y"]), k + " y"
        g = g0.copy()
        o.sweep_y(g, dt, h[1], mat, edge=EDGE_RING)
        assert np.array_equal(g, data[k + "_ycf"]), k + " y_cf"
        g = g0.copy()
        for s in range(10):
            o.run(g, dt, h, mat, 1)
            if s in (0, 1, 9):
                assert np.array_equal(g, data[f"{k}_step{s + 1}"]), f"{k} step {s + 1}"
        p = np.ascontiguousarray(np.moveaxis(g0, -1, 0))
        for s in range(10):
            o.sweep_x(p, dt, h[0], mat)
            o.sweep_y(p, dt, h[1], mat, edge=EDGE_RING)
        assert np.array_equal(np.moveaxis(p, 0, -1), data[k + "_soa10"]), k + " soa"


def test_config0_golden():
    meta, samples = gd.config0()
    nx, ny = meta["nx"], meta["ny"]
    g = np.zeros((ny, nx, 5))
    b = meta["ball"]
    hits = OD.set_ball(g, meta["h"], b["c"], b["r"], b["s"])
    assert hits == b["nodes"] == 305
    assert gd.sha256(g) == b["init_sha256"]
    OD.run(g, meta["dt"], meta["h"], meta["mat"], meta["steps"])
    assert gd.sha256(g) == meta["state_sha256"]
    assert np.array_equal(g.reshape(-1, 5)[samples["index"]], samples["values"])
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "c0.vtk")
        assert OD.save_vtk(g, meta["h"], p) == 0
        import hashlib
        assert hashlib.sha256(open(p, "rb").read()).hexdigest() == meta["vtk_sha256"]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_materials_golden(dtype):
    o = OD if dtype == np.float64 else OF
    m = gd.materials()
    names = ("c1", "c2", "rho", "irhoc1", "irhoc2", "rhoc1", "rhoc2", "rhoc3")
    for inp, exp in zip(m["inputs"], m[np.dtype(dtype).name]):
        got = o.material(*[gd.fromhex(x) for x in inp])
        assert [getattr(got, n) for n in names] == [gd.fromhex(x) for x in exp]


def test_ball_golden():
    for c in gd.balls():
        dt = np.float64 if c["dtype"] == "float64" else np.float32
        o = OD if dt == np.float64 else OF
        nx, ny = c["n"]
        g = np.zeros((ny, nx, 5), dtype=dt)
        o.set_ball(g, c["h"], c["c"], c["r"], c["s"])
        assert gd.sha256(g) == c["sha256"]
        assert np.argwhere(g[..., 2] != 0).tolist() == c["nodes"]


def test_vtk_golden(golden_dir):
    state = np.load(os.path.join(golden_dir, "tiny_vtk_d_state.npy"))
    expect = open(os.path.join(golden_dir, "tiny_vtk_d.vtk"), "rb").read()
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "t.vtk")
        OD.save_vtk(np.ascontiguousarray(state), (0.5, 2.25), p)
        assert open(p, "rb").read() == expect


# --------------------------------------------------------------------------
# 2. closed forms and invariants


def test_material_closed_form():
    # c1=2, c2=1, rho=1: mu=1, lambda=2, rhoc3 = 2*2/(2+2) = 1
    m = OD.material(2.0, 1.0, 1.0)
    assert (m.irhoc1, m.irhoc2, m.rhoc1, m.rhoc2, m.rhoc3) == (0.5, 1.0, 2.0, 1.0, 1.0)
    # rhoc3 = rho*c1*lambda/(lambda+2mu) = rho*(c1^2 - 2 c2^2)/c1
    for c1, c2, rho in [(3.0, 1.0, 2.0), (2.5, 1.25, 0.5)]:
        m = OD.material(c1, c2, rho)
        assert m.rhoc3 == pytest.approx(rho * (c1 * c1 - 2 * c2 * c2) / c1, rel=1e-15)


def test_ball_lattice_count():
    # nodes with integer (i-256)^2 + (j-256)^2 < 100: the Gauss circle count
    count = sum(1 for a in range(-10, 11) for b in range(-10, 11) if a * a + b * b < 100)
    assert count == 305
    g = np.zeros((512, 512, 5))
    assert OD.set_ball(g, (1.0, 1.0), (256.0, 256.0), 10.0, 1.0) == count
    assert g[..., 2].sum() == g[..., 4].sum() == count
    assert not g[..., (0, 1, 3)].any()


def test_superbee_textbook():
    # superbee phi(r) = max(0, min(1, 2r), min(2, r))
    table = {-3.0: 0.0, -0.5: 0.0, 0.0: 0.0, 0.25: 0.5, 0.5: 1.0, 0.75: 1.0, 1.0: 1.0,
             1.5: 1.5, 2.0: 2.0, 3.0: 2.0, 1e9: 2.0}
    for r, v in table.items():
        assert OD.superbee(r) == v
    assert OD.superbee(math.inf) == 2.0
    # the reference's ternary min/max maps NaN (0/0 ratios) to the cap 2
    assert OD.superbee(math.nan) == 2.0


def test_tvd2_special_cases():
    rng = np.random.default_rng(3)
    for u in rng.standard_normal(50):
        # constant data is a fixed point: both ratios are eps/eps = 1
        for c in (0.1, 0.5, 1.0):
            assert OD.tvd2(c, u, u, u, u) == u
    for a, b, u, d in rng.standard_normal((50, 4)):
        # Courant number 1: exact shift of the upwind value, up to one rounding
        got = OD.tvd2(1.0, a, b, u, d)
        assert abs(got - b) <= 2 * np.spacing(max(abs(u), abs(b)))
    # opposite-sign slopes on both faces: phi = 0, first-order upwind
    c = 0.3
    a, b, u, d = 0.0, 1.0, 0.0, 1.0
    assert OD.tvd2(c, a, b, u, d) == u - c * (u - b)


def test_y_independent_data_makes_y_sweep_a_no_op():
    rng = np.random.default_rng(4)
    row = rng.standard_normal((1, 37, 5))
    g = np.ascontiguousarray(np.repeat(row, 11, axis=0))
    g2 = g.copy()
    OD.sweep_y(g2, 0.25, 1.0, (2.0, 1.0, 1.0))
    assert np.array_equal(g, g2)


def _scalar_tvd_line(w, k, sign):
    """1-D superbee TVD advection of one invariant with clamped ends, written
    as a face-flux update (not the oracle's per-node tvd2)."""
    n = len(w)
    p = np.concatenate([[w[0], w[0]], w, [w[-1], w[-1]]])
    if sign < 0:
        p = p[::-1]
    up, u0, dn = p[:-3], p[1:-2], p[2:-1]
    r = ((u0 - up) + 1e-6) / ((dn - u0) + 1e-6)
    phi = np.maximum(0.0, np.maximum(np.minimum(1.0, 2.0 * r), np.minimum(2.0, r)))
    f = u0 + phi * 0.5 * (1.0 - k) * (dn - u0)  # face between u0 and dn
    out = p[2:-2] - k * (f[1:] - f[:-1])
    out = out if sign > 0 else out[::-1]
    assert len(out) == n
    return out


def test_x_sweep_is_four_scalar_advections():
    """Brute force: project on the invariants, advect each with the face-flux
    form, project back.  Matches the oracle to rounding."""
    rng = np.random.default_rng(5)
    nx = 64
    g = np.zeros((3, nx, 5))
    g[:, 20:40, :] = rng.standard_normal((1, 20, 5))
    m = OD.material(2.0, 1.0, 1.0)
    dt, hx = 0.25, 1.0
    k1, k2 = dt * m.c1 / hx, dt * m.c2 / hx
    vx, vy, sxx, sxy, syy = (g[0, :, q] for q in range(5))
    w1 = vx - sxx * m.irhoc1
    w2 = vx + sxx * m.irhoc1
    w3 = vy - sxy * m.irhoc2
    w4 = vy + sxy * m.irhoc2
    n1 = _scalar_tvd_line(w1, k1, +1) - w1
    n2 = _scalar_tvd_line(w2, k1, -1) - w2
    n3 = _scalar_tvd_line(w3, k2, +1) - w3
    n4 = _scalar_tvd_line(w4, k2, -1) - w4
    e1, e2, e3, e4 = 0.5 * n1, 0.5 * n2, 0.5 * n3, 0.5 * n4
    expect = np.stack([vx + e1 + e2, vy + e3 + e4, sxx + (e2 - e1) * m.rhoc1,
                       sxy + m.rhoc2 * (e4 - e3), syy + (e2 - e1) * m.rhoc3], axis=-1)
    got = g.copy()
    OD.sweep_x(got, dt, hx, m)
    np.testing.assert_allclose(got[0], expect, rtol=0, atol=1e-13)
    # rows are independent
    assert np.array_equal(got[0], got[1]) and np.array_equal(got[1], got[2])


@pytest.mark.parametrize("mat", [(2.0, 1.0, 1.5), (3.1, 1.6, 2.3)])
def test_sweeps_keep_their_null_space_combination(mat):
    """An x-sweep adds (e2-e1)*rhoc1 to sxx and (e2-e1)*rhoc3 to syy, so
    syy - (rhoc3/rhoc1)*sxx does not change; a y-sweep keeps
    sxx - (rhoc3/rhoc1)*syy.  A swapped rhoc1/rhoc3 or stress field in the
    increments breaks this."""
    m = OD.material(*mat)
    q = m.rhoc3 / m.rhoc1
    rng = np.random.default_rng(9)
    g = rng.standard_normal((20, 24, 5))  # (vx, vy, sxx, sxy, syy)
    x_keep = lambda a: a[..., 4] - q * a[..., 2]  # noqa: E731
    y_keep = lambda a: a[..., 2] - q * a[..., 4]  # noqa: E731
    for sweep, keep, other in ((OD.sweep_x, x_keep, y_keep), (OD.sweep_y, y_keep, x_keep)):
        a = g.copy()
        sweep(a, 0.2, 1.0, mat)
        np.testing.assert_allclose(keep(a), keep(g), rtol=0, atol=1e-12)
        # the stresses did change, and the other sweep's combination with them
        assert np.abs(a[..., 2] - g[..., 2]).max() > 1e-2
        assert np.abs(other(a) - other(g)).max() > 1e-2


def test_conservation_while_edges_are_quiet():
    g = np.zeros((96, 96, 5))
    OD.set_ball(g, (1.0, 1.0), (48.0, 48.0), 8.0, 1.0)
    s0 = g.reshape(-1, 5).sum(axis=0)
    OD.run(g, 0.25, (1.0, 1.0), (2.0, 1.0, 1.0), 20)
    s1 = g.reshape(-1, 5).sum(axis=0)
    # edges still quiet (fronts travel 20*0.25*2 = 10 cells)
    assert not g[:2].any() and not g[-2:].any() and not g[:, :2].any() and not g[:, -2:].any()
    np.testing.assert_allclose(s1, s0, rtol=0, atol=1e-10)


def test_clamp_edge_equals_two_ghost_nodes():
    rng = np.random.default_rng(6)
    g = rng.standard_normal((9, 17, 5))
    padded = np.ascontiguousarray(np.concatenate([g[:, :1], g[:, :1], g, g[:, -1:], g[:, -1:]],
                                                 axis=1))
    a = g.copy()
    OD.sweep_x(a, 0.2, 1.0, (2.0, 1.0, 1.5))
    OD.sweep_x(padded, 0.2, 1.0, (2.0, 1.0, 1.5))
    assert np.array_equal(a, padded[:, 2:-2])


def test_x_and_y_sweeps_are_transposes():
    """The y-sweep of a grid is the x-sweep of its transpose with (vx, sxx)
    and (vy, syy) exchanged."""
    rng = np.random.default_rng(7)
    g = rng.standard_normal((13, 21, 5))
    a = g.copy()
    OD.sweep_y(a, 0.2, 1.0, (2.0, 1.0, 1.5))
    perm = [1, 0, 4, 3, 2]  # vx<->vy, sxx<->syy
    t = np.ascontiguousarray(g.transpose(1, 0, 2)[..., perm])
    OD.sweep_x(t, 0.2, 1.0, (2.0, 1.0, 1.5))
    assert np.array_equal(a, t[..., perm].transpose(1, 0, 2))


def test_ring_edge_differs_from_clamp_only_on_last_two_rows():
    rng = np.random.default_rng(8)
    g = rng.standard_normal((12, 10, 5))
    a, b = g.copy(), g.copy()
    OD.sweep_y(a, 0.2, 1.0, (2.0, 1.0, 1.5), edge=EDGE_CLAMP)
    OD.sweep_y(b, 0.2, 1.0, (2.0, 1.0, 1.5), edge=EDGE_RING)
    assert np.array_equal(a[:-2], b[:-2])
    assert not np.array_equal(a[-2:], b[-2:])
