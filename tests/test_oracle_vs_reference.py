"""The oracle against the reference's own simple/ path, compiled from the
reference sources where they lie (oracle/_ref, `make -C oracle ref`; skipped
where it is not built), on inputs drawn afresh: random shapes down to three
nodes a side, Courant numbers from 0 to 1, random materials and spacings,
Gaussian fields and fields built from a few exact values (slopes of exactly
-eps, so zero and infinite ratios and 0/0), every sweep entry point.  The
golden fixtures pin the oracle on fixed cases; this pins it on cases nobody
chose.  Both are uncontracted C on the host, so FP64 and FP32 agree bit for
bit (uint64/uint32 views: signed zeros and NaN patterns included)."""
import ctypes

import numpy as np
import pytest

from oracle import EDGE_RING, Oracle, Reference

SPECIAL = np.array([-0.0, 0.0, 1e-6, -1e-6, 2e-6, 0.5, -1.0, 1.0])
DTYPES = [np.float64, np.float32]


def _skip_unless_built(dtype):
    if not Reference.available(dtype):
        pytest.skip("oracle/_ref not built (the reference sources are not present)")


def _case(seed, dtype):
    rng = np.random.default_rng(seed)
    nx, ny = (int(v) for v in rng.integers(3, 200 if seed % 8 == 7 else 48, size=2))
    c1 = float(rng.uniform(1.0, 3.0))
    mat = (c1, c1 / float(rng.uniform(1.5, 2.2)), float(rng.uniform(1.0, 3.0)))
    h = (float(rng.uniform(0.5, 2.0)), float(rng.uniform(0.5, 2.0)))
    # Courant numbers c1 dt / h in [0, 1] along both axes, both ends included
    frac = (0.0, 1.0)[(seed // 10) % 2] if seed % 10 == 5 else float(rng.uniform(0.0, 1.0))
    dt = frac * min(h) / c1
    if seed % 3 == 0:
        g = SPECIAL[rng.integers(0, len(SPECIAL), (ny, nx, 5))]
    else:
        g = rng.standard_normal((ny, nx, 5)) * 10.0 ** float(rng.uniform(-3, 3))
    return nx, ny, dt, h, mat, np.ascontiguousarray(g.astype(dtype))


def _bits(a):
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def _same(a, b):
    return np.array_equal(_bits(a), _bits(b))


@pytest.mark.parametrize("dtype", DTYPES)
def test_random_cases_every_entry_point(dtype):
    _skip_unless_built(dtype)
    ref, orc = Reference(dtype), Oracle(dtype)
    for seed in range(80):
        nx, ny, dt, h, mat, g0 = _case(seed, dtype)
        steps = 1 + seed % 7

        def fresh(stype=1):
            t = ref.task(nx, ny, dt, h, mat, stype=stype)
            ref.view(t)[...] = g0 if stype == 1 else np.moveaxis(g0, -1, 0)
            return t

        # single x- and y-sweeps
        for fn, axis in ((ref.lib.le_step_x, 0), (ref.lib.le_step_y, 1)):
            t = fresh()
            fn(ctypes.byref(t))
            want = ref.view(t).copy()
            ref.lib.le_free_task(ctypes.byref(t))
            got = g0.copy()
            if axis == 0:
                orc.sweep_x(got, dt, h[0], mat)
            else:
                orc.sweep_y(got, dt, h[1], mat)
            assert _same(got, want), (seed, axis)
        # le_step, several steps
        t = fresh()
        for _ in range(steps):
            ref.lib.le_step(ctypes.byref(t))
        want = ref.view(t).copy()
        ref.lib.le_free_task(ctypes.byref(t))
        got = g0.copy()
        orc.run(got, dt, h, mat, steps)
        assert _same(got, want), (seed, "le_step", steps)
        # le_step_soa: SOA planes, the y-sweep's ring edge
        t = fresh(stype=0)
        for _ in range(steps):
            ref.lib.le_step_soa(ctypes.byref(t))
        want = ref.view(t).copy()
        ref.lib.le_free_task(ctypes.byref(t))
        got = np.ascontiguousarray(np.moveaxis(g0, -1, 0))
        for _ in range(steps):
            orc.sweep_x(got, dt, h[0], mat, soa=True)
            orc.sweep_y(got, dt, h[1], mat, edge=EDGE_RING, soa=True)
        assert _same(got, want), (seed, "le_step_soa", steps)


@pytest.mark.parametrize("dtype", DTYPES)
def test_random_balls(dtype):
    """le_set_ball: random centres (inside and outside the grid), radii and
    spacings; the lattice test is a strict inequality on a libm sqrt."""
    _skip_unless_built(dtype)
    ref, orc = Reference(dtype), Oracle(dtype)
    rng = np.random.default_rng(99)
    for _ in range(30):
        nx, ny = (int(v) for v in rng.integers(3, 80, size=2))
        h = (float(rng.uniform(0.3, 2.0)), float(rng.uniform(0.3, 2.0)))
        c = (float(rng.uniform(-10, nx * h[0] + 10)), float(rng.uniform(-10, ny * h[1] + 10)))
        r, s = float(rng.uniform(0.0, 30.0)), float(rng.normal())
        t = ref.task(nx, ny, 0.1, h, (2.0, 1.0, 1.0))
        ref.lib.le_set_ball(ctypes.byref(t), ref.abi.le_vec2(*c), r, s)
        want = ref.view(t).copy()
        ref.lib.le_free_task(ctypes.byref(t))
        got = np.zeros((ny, nx, 5), dtype=dtype)
        orc.set_ball(got, h, c, r, s)
        assert _same(got, want), (nx, ny, h, c, r, s)
