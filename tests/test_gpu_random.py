"""The HIP library against the oracle on cases drawn afresh: random shapes from
3 to 300 nodes a side, Courant numbers from 0 to 1, random materials and
spacings, Gaussian fields spanning six decades and fields of exact special
values (slopes of exactly -eps: zero, infinite and 0/0 ratios), each case
through every step path at a random launch shape: paired le_step calls,
le_step_n's two-step passes (random segment height, window width and, at 64
lanes, one or two wavefronts), the single x- and y-sweeps and the SOA ring
edge.  FP64 bit for bit (uint64 views); FP32 on the Gaussian fields within the
tier's 1e-5 relative L2 of the FP64 oracle."""
import numpy as np
import pytest

import le2d_amd
from le2d_amd import ST_SOA
from oracle import EDGE_RING, Oracle

pytestmark = pytest.mark.gpu
OD = Oracle(np.float64)
SPECIAL = np.array([-0.0, 0.0, 1e-6, -1e-6, 2e-6, 0.5, -1.0, 1.0])


def _case(seed):
    rng = np.random.default_rng(1000 + seed)
    nx, ny = (int(v) for v in rng.integers(3, 300 if seed % 4 == 3 else 64, size=2))
    c1 = float(rng.uniform(1.0, 3.0))
    mat = (c1, c1 / float(rng.uniform(1.5, 2.2)), float(rng.uniform(1.0, 3.0)))
    h = (float(rng.uniform(0.5, 2.0)), float(rng.uniform(0.5, 2.0)))
    dt = float(rng.uniform(0.0, 1.0)) * min(h) / c1
    special = seed % 3 == 0
    if special:
        g = SPECIAL[rng.integers(0, len(SPECIAL), (ny, nx, 5))]
    else:
        g = rng.standard_normal((ny, nx, 5)) * 10.0 ** float(rng.uniform(-3, 3))
    steps = int(rng.integers(1, 8))
    lanes = int(rng.choice([64, 128, 256]))
    shape = (int(rng.integers(1, 70)), lanes, int(rng.integers(0, 2)) if lanes == 64 else 0)
    return nx, ny, dt, h, mat, np.ascontiguousarray(g), steps, shape, special


def _same(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("block", range(4))
def test_random_cases_fp64(block):
    for seed in range(12 * block, 12 * block + 12):
        nx, ny, dt, h, mat, g0, steps, shape, _ = _case(seed)
        ref = g0.copy()
        OD.run(ref, dt, h, mat, steps)
        # paired le_step calls at the tuned default shape
        with le2d_amd.Task(nx, ny, dt, h, mat) as t:
            t.load(g0)
            for _ in range(steps):
                t.step()
            assert _same(t.grid(), ref), (seed, "le_step", steps)
        # le_step_n at a random two-step launch shape
        with le2d_amd.Task(nx, ny, dt, h, mat) as t:
            t.load(g0)
            t.set_step_pairs(shape[0])
            t.set_step_pair_window(shape[1])
            t.set_step_pair_split(shape[2])
            t.step_n(steps)
            assert _same(t.grid(), ref), (seed, "le_step_n", steps, shape)
        # single sweeps
        with le2d_amd.Task(nx, ny, dt, h, mat) as t:
            t.load(g0)
            t.step_x()
            rx = g0.copy()
            OD.sweep_x(rx, dt, h[0], mat)
            assert _same(t.grid(), rx), (seed, "le_step_x")
            t.step_y()
            OD.sweep_y(rx, dt, h[1], mat)
            assert _same(t.grid(), rx), (seed, "le_step_y")
        # SOA planes, the ring edge
        p0 = np.ascontiguousarray(np.moveaxis(g0, -1, 0))
        rs = p0.copy()
        for _ in range(steps):
            OD.sweep_x(rs, dt, h[0], mat, soa=True)
            OD.sweep_y(rs, dt, h[1], mat, edge=EDGE_RING, soa=True)
        with le2d_amd.Task(nx, ny, dt, h, mat, stype=ST_SOA) as t:
            t.load(p0)
            for _ in range(steps):
                t.step()
            assert _same(t.grid(), rs), (seed, "le_step_soa", steps)


def test_random_cases_fp32():
    for seed in range(48):
        nx, ny, dt, h, mat, g0, steps, shape, special = _case(seed)
        if special:
            continue  # exact -eps slopes: FP32 rounds them differently by design
        ref = g0.copy()
        OD.run(ref, dt, h, mat, steps)
        with le2d_amd.Task(nx, ny, dt, h, mat, dtype=np.float32) as t:
            t.load(g0.astype(np.float32))
            t.set_step_pairs(shape[0])
            t.set_step_pair_window(shape[1])
            t.set_step_pair_split(shape[2])
            t.step_n(steps)
            got = t.grid().astype(np.float64)
        rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert rel < 1e-5, (seed, rel, nx, ny, steps, shape)
