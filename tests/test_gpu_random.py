"""The HIP library against the oracle on cases drawn afresh: random shapes from
3 to 300 nodes a side, Courant numbers from 0 to 1, random materials and
spacings, Gaussian fields spanning six decades and fields of exact special
values (slopes of exactly -eps: zero, infinite and 0/0 ratios), each case
through every step path at a random launch shape: paired le_step calls,
le_step_n's two-step passes (random segment height, window width and, at 64
lanes, one or two wavefronts), the single x- and y-sweeps and the SOA ring
edge; then batches of random instances in every launch form, random
material planes, and the minmod and CIR builds.  FP64 bit for bit (uint64 views); FP32 on the Gaussian
fields within the tier's 1e-5 relative L2 of the FP64 oracle."""
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


def test_random_batches_fp64():
    """Batches of 3 to 12 random instances (each its own shape, material,
    spacing, time step and field) in every launch form at a random window
    width and segment height, against the oracle per instance."""
    forms = ["tiles", "persistent", "pairs", "pairs_split"]
    for seed in range(8):
        rng = np.random.default_rng(5000 + seed)
        insts = [_case(100 * seed + k)[:6] for k in range(int(rng.integers(3, 13)))]
        steps = int(rng.integers(1, 10))
        shape = (int(rng.choice([64, 128, 256])), int(rng.integers(1, 80)), forms[seed % 4])
        with le2d_amd.Batch(np.float64) as b:
            for nx, ny, dt, h, mat, _ in insts:
                b.add(nx, ny, dt, h, mat)
            b.commit()
            for k, inst in enumerate(insts):
                b.load(k, inst[5])
            b.set_shape(*shape)
            b.step(steps)
            for k, (nx, ny, dt, h, mat, g0) in enumerate(insts):
                ref = g0.copy()
                OD.run(ref, dt, h, mat, steps)
                assert _same(b.read(k), ref), (seed, k, shape, steps)


def _ids(rng, nx, ny, nm, kind):
    if kind == 0:  # every node its own draw
        return rng.integers(0, nm, size=(ny, nx)).astype(np.uint8)
    if kind == 1:  # bands of random widths along x
        cuts = np.sort(rng.integers(0, nx, size=nm - 1))
        return np.searchsorted(cuts, np.arange(nx), side="right")[None, :].repeat(ny, 0).astype(
            np.uint8)
    y, x = np.mgrid[0:ny, 0:nx]  # blobs
    ids = np.zeros((ny, nx), np.uint8)
    for m in range(1, nm):
        cx, cy, r = rng.uniform(0, nx), rng.uniform(0, ny), rng.uniform(1, max(nx, ny) / 2)
        ids[(x - cx) ** 2 + (y - cy) ** 2 < r * r] = m
    return ids


def test_random_material_planes_fp64():
    """One to three random materials on per-node, band and blob layouts; the
    time step from the fastest material; le_step_n at a random two-step
    launch shape and paired le_step calls, against the oracle's node-local
    scheme."""
    for seed in range(18):
        nx, ny, _, h, _, g0, steps, shape, _ = _case(300 + seed)
        rng = np.random.default_rng(7000 + seed)
        nm = int(rng.integers(1, 4))
        tab = []
        for _ in range(nm):
            c1 = float(rng.uniform(1.0, 3.0))
            tab.append((c1, c1 / float(rng.uniform(1.5, 2.2)), float(rng.uniform(1.0, 3.0))))
        dt = float(rng.uniform(0.0, 1.0)) * min(h) / max(m[0] for m in tab)
        ids = _ids(rng, nx, ny, nm, seed % 3)
        ref = g0.copy()
        OD.run_mat(ref, dt, h, tab, ids, steps)
        for paired_calls in (False, True):
            with le2d_amd.Task(nx, ny, dt, h, tab[0]) as t:
                t.load(g0)
                t.set_materials(tab, ids)
                if paired_calls:
                    for _ in range(steps):
                        t.step()
                else:
                    t.set_step_pairs(shape[0])
                    t.set_step_pair_window(shape[1])
                    t.set_step_pair_split(shape[2])
                    t.step_n(steps)
                assert _same(t.grid(), ref), (seed, nm, seed % 3, steps, shape, paired_calls)


@pytest.mark.parametrize("limiter", ["minmod", "cir"])
def test_random_cases_other_limiters(limiter):
    """The minmod and CIR libraries on 16 of the random cases (paired
    le_step calls, le_step_n at the case's launch shape, the SOA ring edge)
    against the oracle built with the same limiter, bit for bit."""
    o = Oracle(np.float64, limiter)
    for seed in range(16):
        nx, ny, dt, h, mat, g0, steps, shape, _ = _case(seed)
        ref = g0.copy()
        o.run(ref, dt, h, mat, steps)
        with le2d_amd.Task(nx, ny, dt, h, mat, limiter=limiter) as t:
            t.load(g0)
            for _ in range(steps):
                t.step()
            assert _same(t.grid(), ref), (limiter, seed, "le_step", steps)
        with le2d_amd.Task(nx, ny, dt, h, mat, limiter=limiter) as t:
            t.load(g0)
            t.set_step_pairs(shape[0])
            t.set_step_pair_window(shape[1])
            t.set_step_pair_split(shape[2])
            t.step_n(steps)
            assert _same(t.grid(), ref), (limiter, seed, "le_step_n", steps, shape)
        p0 = np.ascontiguousarray(np.moveaxis(g0, -1, 0))
        rs = p0.copy()
        for _ in range(steps):
            o.sweep_x(rs, dt, h[0], mat, soa=True)
            o.sweep_y(rs, dt, h[1], mat, edge=EDGE_RING, soa=True)
        with le2d_amd.Task(nx, ny, dt, h, mat, stype=ST_SOA, limiter=limiter) as t:
            t.load(p0)
            for _ in range(steps):
                t.step()
            assert _same(t.grid(), rs), (limiter, seed, "le_step_soa", steps)
