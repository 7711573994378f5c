"""FP64 results compared as bit patterns (signed zeros included), on inputs
built from a few exact values: +-0, +-1e-6, 2e-6, 0.5, +-1.  Differences of
these hit the limiter's edge cases exactly -- slopes of exactly -eps, so
zero numerators, ratios of +-0 and infinite ratios -- which random inputs
never reach.  The FP64 superbee gives +0 for a ratio of -0, where the
reference's ternaries give -0; these grids reach such ratios, and the states
must still agree bit for bit.  np.array_equal would treat -0 and +0 as
equal, so the uint64 views are compared.  Each limiter build (superbee, minmod, CIR) against the
oracle built with the same limiter; then magnitudes at both ends of the FP64
range, and non-finite inputs."""
import numpy as np
import pytest

import le2d_amd
from oracle import Oracle

pytestmark = pytest.mark.gpu
SPECIAL = np.array([-0.0, 0.0, 1e-6, -1e-6, 2e-6, 0.5, -1.0, 1.0])
ARGS = (0.25, (1.0, 1.0), (2.0, 1.0, 1.0))


def _grid(seed, ny, nx, stresses):
    rng = np.random.default_rng(seed)
    g = SPECIAL[rng.integers(0, len(SPECIAL), (ny, nx, 5))]
    if not stresses:
        g[..., 2:] = 0.0  # invariants then equal the velocities exactly
    return g


def _negzero_ratio_faces(v):
    """Faces of the +c invariant along the last axis where the slope ratio
    is (+0) / (negative) = -0: (u0 - um) + eps == 0 and (u1 - u0) + eps < 0."""
    num = (v[..., 1:-1] - v[..., :-2]) + 1e-6
    den = (v[..., 2:] - v[..., 1:-1]) + 1e-6
    return int(np.count_nonzero((num == 0) & (den < 0)))


def _same_bits(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("stresses", [False, True])
@pytest.mark.parametrize("pairs", [0, 16])
@pytest.mark.parametrize("limiter", ["superbee", "minmod", "cir"])
def test_special_values_bit_identical(stresses, pairs, limiter):
    ny, nx = 96, 300
    g0 = _grid(17 + stresses, ny, nx, stresses)
    if not stresses:
        # the crafted input really reaches r = -0 in both sweeps
        assert _negzero_ratio_faces(g0[..., 0]) > 0
        assert _negzero_ratio_faces(np.swapaxes(g0[..., 1], 0, 1)) > 0
    for steps in (1, 4):
        with le2d_amd.Task(nx, ny, *ARGS, limiter=limiter) as t:
            t.load(g0)
            t.set_step_pairs(pairs)
            t.step_n(steps)
            got = t.grid().copy()
        ref = g0.copy()
        Oracle(np.float64, limiter).run(ref, *ARGS, steps)
        assert _same_bits(got, ref), (steps, int(np.count_nonzero(got.view(np.uint64) != ref.view(np.uint64))))


def _extreme_cases():
    rng = np.random.default_rng(3)
    g = rng.standard_normal((60, 70, 5))
    m = g.copy()
    m[:20] *= 1e280
    m[20:40] *= 1e-290
    h = g.copy()
    h[::7, ::5, 2] = 1e307
    return {"huge": g * 1e200, "subnormal": g * 1e-300, "mixed": m, "near_max": h}


@pytest.mark.parametrize("name", ["huge", "subnormal", "mixed", "near_max"])
@pytest.mark.parametrize("pairs", [0, 16])
def test_extreme_magnitudes_bit_identical(name, pairs):
    """Values near the top of the FP64 range, subnormal ones and both in one
    grid: the slope ratios then need the division's operand scaling, and
    subnormal operands stay subnormal; bit for bit against the oracle."""
    g0 = _extreme_cases()[name]
    with le2d_amd.Task(70, 60, *ARGS) as t:
        t.load(g0)
        t.set_step_pairs(pairs)
        t.step_n(4)
        got = t.grid().copy()
    ref = g0.copy()
    Oracle(np.float64).run(ref, *ARGS, 4)
    assert _same_bits(got, ref)


@pytest.mark.parametrize("pairs", [0, 16])
def test_non_finite_inputs_spread_as_in_the_reference(pairs):
    """A NaN and two infinities in the input: after four steps the NaNs sit
    on exactly the nodes where the oracle has them, and every other value is
    bit-identical.  The NaNs' own bit patterns (sign, payload) are not
    compared: the reference's x86 arithmetic propagates an operand's NaN,
    the GPU returns its own quiet NaN."""
    g0 = np.random.default_rng(4).standard_normal((60, 70, 5))
    g0[10, 10, 0] = np.nan
    g0[30, 40, 3] = np.inf
    g0[50, 5, 4] = -np.inf
    with le2d_amd.Task(70, 60, *ARGS) as t:
        t.load(g0)
        t.set_step_pairs(pairs)
        t.step_n(4)
        got = t.grid().copy()
    ref = g0.copy()
    Oracle(np.float64).run(ref, *ARGS, 4)
    nan = np.isnan(ref)
    assert nan.any() and np.array_equal(np.isnan(got), nan)
    assert _same_bits(got[~nan], ref[~nan])
