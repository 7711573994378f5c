"""FP64 results compared as bit patterns (signed zeros included), on inputs
built from a few exact values: +-0, +-1e-6, 2e-6, 0.5, +-1.  Differences of
these hit the limiter's edge cases exactly -- slopes of exactly -eps, so
zero numerators, ratios of +-0 and infinite ratios -- which random inputs
never reach.  np.array_equal would treat -0 and +0 as equal, so the uint64
views are compared.  Each limiter build (superbee, minmod, CIR) against the
oracle built with the same limiter."""
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
