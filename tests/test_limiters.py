"""The reference's other two flux limiters (simple/le_core.c:36-43): minmod and
CIR, a build-time choice there (`#define limiter limiter_superbee`) as here
(oracle: ORC_LIMITER; library: make LIMITER=minmod|cir, `make limiters`).

The reference's sources cannot build these variants without editing them, so
the oracle's minmod and CIR builds are pinned by the mathematics:
  * the limiter values of the reference's ternary definitions;
  * CIR (phi = 0) is first-order upwind, whose transport of a one-node pulse
    in one invariant is the binomial distribution;
  * minmod is Lax-Wendroff where both faces see r >= 1 (concave increasing
    data) and, up to eps, Beam-Warming where 0 < r < 1 (convex increasing);
  * both keep each advected invariant's total variation from growing.
The HIP libraries of the two limiters are checked against these oracles bit
for bit on the GPU (test_gpu_limiters.py); here, that they load and report
the limiter they were built with.
"""
import ctypes
import math
import os

import numpy as np
import pytest

from oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OM = Oracle(np.float64, "minmod")
OC = Oracle(np.float64, "cir")
OS = Oracle(np.float64)


def test_oracle_builds_report_their_limiter():
    assert OS.lib.orc_limiter_kind() == 0
    assert OM.lib.orc_limiter_kind() == 1
    assert OC.lib.orc_limiter_kind() == 2
    with pytest.raises(ValueError):
        Oracle(np.float64, "vanleer")


def test_limiter_values():
    # minmod phi(r) = max(0, min(1, r)) with the reference's ternaries
    # (le_min(a, b) = a > b ? b : a, le_max(a, b) = a > b ? a : b)
    table = {-3.0: 0.0, -0.5: 0.0, 0.0: 0.0, 0.25: 0.25, 0.5: 0.5, 0.999: 0.999, 1.0: 1.0,
             1.5: 1.0, 2.0: 1.0, 1e9: 1.0}
    for o in (OM, Oracle(np.float32, "minmod")):
        for r, v in table.items():
            assert o.limiter(r) == float(o.dtype.type(v))
        assert o.limiter(math.inf) == 1.0
        assert o.limiter(-math.inf) == 0.0
        # min(1, NaN) keeps the 1, and max(0, -0) keeps the -0
        assert o.limiter(math.nan) == 1.0
        assert math.copysign(1.0, o.limiter(-0.0)) == -1.0
    for r in (-2.0, -0.0, 0.0, 0.5, 1.0, 7.0, math.inf, math.nan):
        v = OC.limiter(r)
        assert v == 0.0 and math.copysign(1.0, v) == 1.0
    # the superbee build's limiter is the superbee function
    for r in (-1.0, 0.3, 0.7, 1.4, 5.0, math.nan):
        assert OS.limiter(r) == OS.superbee(r)


def _line(nx, ny=4):
    return np.zeros((ny, nx, 5))


def _inv(o, g, mat, which):
    """x-sweep invariant along the first row: 'p+' = vx - sxx/(rho c1) (moves
    towards +x), 's+' = vy - sxy/(rho c2)."""
    m = o.material(*mat)
    row = g[0]
    if which == "p+":
        return row[:, 0] - row[:, 2] * m.irhoc1
    return row[:, 1] - row[:, 3] * m.irhoc2


@pytest.mark.parametrize("wave", ["p+", "s+"])
def test_cir_moves_a_pulse_as_the_binomial_distribution(wave):
    """phi = 0: each step w_i <- (1-k) w_i + k w_{i-1} for the right-moving
    invariant, so a unit pulse at i0 becomes C(n, j) k^j (1-k)^(n-j) at
    i0 + j after n steps (data uniform in y: the y-sweep is a no-op)."""
    mat = (2.0, 1.2, 1.5)
    m = OC.material(*mat)
    dt, h, n, i0, nx = 0.15, (1.0, 1.0), 40, 20, 120
    g = _line(nx)
    if wave == "p+":
        g[:, i0, 0] = 0.5
        g[:, i0, 2] = -0.5 * m.rhoc1  # w = 1, the left-moving partner 0
        k = dt * m.c1 / h[0]
    else:
        g[:, i0, 1] = 0.5
        g[:, i0, 3] = -0.5 * m.rhoc2
        k = dt * m.c2 / h[0]
    OC.run(g, dt, h, mat, n)
    w = _inv(OC, g, mat, wave)
    expect = np.zeros(nx)
    for j in range(n + 1):
        expect[i0 + j] = math.comb(n, j) * k**j * (1 - k) ** (n - j)
    np.testing.assert_allclose(w, expect, rtol=0, atol=1e-13)
    # superbee does not spread it that way (it steepens the pulse)
    g2 = _line(nx)
    if wave == "p+":
        g2[:, i0, 0], g2[:, i0, 2] = 0.5, -0.5 * m.rhoc1
    else:
        g2[:, i0, 1], g2[:, i0, 3] = 0.5, -0.5 * m.rhoc2
    OS.run(g2, dt, h, mat, n)
    assert np.abs(_inv(OS, g2, mat, wave) - expect).max() > 1e-2


def _one_x_step(o, w, mat, dt):
    """One step (h = 1) of the field whose right-moving P invariant is w (and
    whose other invariants are 0); returns the new invariant."""
    m = o.material(*mat)
    g = _line(len(w))
    g[:, :, 0] = 0.5 * w
    g[:, :, 2] = -0.5 * w * m.rhoc1
    o.run(g, dt, (1.0, 1.0), mat, 1)
    return _inv(o, g, mat, "p+")


def test_minmod_is_lax_wendroff_on_concave_increasing_data():
    """r = (du_left + eps)/(du_right + eps) >= 1 on every face of a concave
    increasing profile, where minmod's phi = 1: the Lax-Wendroff update
    u - k/2 (u+ - u-) + k^2/2 (u+ - 2u + u-)."""
    mat = (2.0, 1.0, 1.0)
    dt = 0.2
    k = dt * 2.0
    i = np.arange(80, dtype=np.float64)
    w = np.sqrt(1.0 + i)
    got = _one_x_step(OM, w, mat, dt)
    um, u, up = w[:-2], w[1:-1], w[2:]
    lw = u - 0.5 * k * (up - um) + 0.5 * k * k * (up - 2 * u + um)
    np.testing.assert_allclose(got[3:-3], lw[2:-2], rtol=0, atol=1e-13)
    # superbee's phi = r there: not Lax-Wendroff
    sb = _one_x_step(OS, w, mat, dt)
    assert np.abs(sb[3:-3] - lw[2:-2]).max() > 1e-4


def test_minmod_is_beam_warming_on_convex_increasing_data():
    """0 < r < 1 on every face of a convex increasing profile: phi = r, so the
    face flux is u + (1-k)/2 (u - u_prev) up to terms of order eps (second
    order upwind, Beam-Warming)."""
    mat = (2.0, 1.0, 1.0)
    dt = 0.2
    k = dt * 2.0
    i = np.arange(80, dtype=np.float64)
    w = (i / 8.0) ** 2
    got = _one_x_step(OM, w, mat, dt)
    umm, um, u = w[:-2], w[1:-1], w[2:]
    bw = u - k * (u - um) - 0.5 * k * (1 - k) * (u - 2 * um + umm)
    np.testing.assert_allclose(got[4:-3], bw[2:-3], rtol=0, atol=2e-6)


@pytest.mark.parametrize("o", [OM, OC, OS], ids=["minmod", "cir", "superbee"])
def test_limited_schemes_do_not_raise_total_variation(o):
    """Each of the three limiters is TVD for Courant numbers up to 1: the
    advected invariant's total variation does not grow."""
    rng = np.random.default_rng(17)
    mat = (2.0, 1.0, 1.0)
    for trial in range(5):
        w = np.cumsum(rng.standard_normal(60)) * (trial + 1)
        w[:5] = w[5]
        w[-5:] = w[-6]
        tv0 = np.abs(np.diff(w)).sum()
        got = _one_x_step(o, w, mat, 0.45)
        assert np.abs(np.diff(got)).sum() <= tv0 * (1 + 1e-12)


@pytest.mark.parametrize("name,code", [("minmod", 1), ("cir", 2), ("superbee", 0)])
def test_hip_limiter_libraries_report_their_limiter(name, code):
    """Loads each build of the HIP library (no GPU call) and asks which
    limiter it was built with."""
    d = "lib" if name == "superbee" else "lib_" + name
    for suffix in ("d", "f"):
        path = os.path.join(ROOT, "le2d_amd", d, f"liblecore_{suffix}.so")
        if not os.path.exists(path):
            pytest.skip(f"{path} not built (make -C le2d_amd/csrc all limiters)")
        lib = ctypes.CDLL(path)
        lib.le2d_limiter.restype = ctypes.c_int
        assert lib.le2d_limiter() == code
