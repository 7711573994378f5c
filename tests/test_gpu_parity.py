"""GPU parity of the FP64 sweeps through the C ABI (liblecore_d.so).

Bar: bit-identical to the reference (golden vectors from the compiled
simple/ code) and to the oracle on seeded inputs.  The library is built with
-ffp-contract=off, so no tolerance is allowed for FP64.
"""
import hashlib
import os
import tempfile

import numpy as np
import pytest

import golden_data as gd
import le2d_amd
from le2d_amd import ST_AOS, ST_SOA, Task
from oracle import EDGE_RING, Oracle

pytestmark = pytest.mark.gpu
OD = Oracle(np.float64)


def make(g0, dt, h, mat, stype=ST_AOS):
    ny, nx = (g0.shape[0], g0.shape[1]) if stype == ST_AOS else (g0.shape[1], g0.shape[2])
    t = Task(nx, ny, dt, h, mat, stype=stype)
    t.load(g0)
    return t


def test_tiny_golden_all_entry_points():
    cases, data = gd.tiny(np.float64)
    for c in cases:
        k, dt, h, mat = c["key"], c["dt"], c["h"], c["mat"]
        g0 = data[k + "_init"]
        with make(g0, dt, h, mat) as t:
            t.step_x()
            assert gd.same_bits(t.grid(), data[k + "_x"]), k + " le_step_x"
        with make(g0, dt, h, mat) as t:
            t.step_y()
            assert gd.same_bits(t.grid(), data[k + "_y"]), k + " le_step_y"
        with make(g0, dt, h, mat) as t:
            t.step_y_cf(max(1, c["nx"] // 2))
            assert gd.same_bits(t.grid(), data[k + "_ycf"]), k + " le_step_y_cf"
        with make(g0, dt, h, mat) as t:
            for s in range(10):
                t.step()
                if s in (0, 1, 9):
                    assert gd.same_bits(t.grid(), data[f"{k}_step{s + 1}"]), f"{k} step {s + 1}"
        with make(np.ascontiguousarray(np.moveaxis(g0, -1, 0)), dt, h, mat, ST_SOA) as t:
            for s in range(10):
                t.step()
            assert gd.same_bits(np.moveaxis(t.grid(), 0, -1), data[k + "_soa10"]), k + " soa"


def test_fused_step_equals_separate_sweeps():
    """le_step (one fused kernel) and le_step_x + le_step_y (two kernels) agree
    bit for bit, on both edge modes and with segments of 1 row."""
    rng = np.random.default_rng(15)
    g0 = rng.standard_normal((131, 250, 5))
    for stype in (ST_AOS, ST_SOA):
        init = g0 if stype == ST_AOS else np.ascontiguousarray(np.moveaxis(g0, -1, 0))
        res = []
        for fused in (True, False):
            with make(init, 0.3, (1.0, 0.8), (2.0, 1.0, 1.5), stype) as t:
                t.set_tiles(xy_seg=1 if fused else 0, xy_block=64 if fused else 0)
                for _ in range(3):
                    if fused:
                        t.step()
                    else:
                        t.step_x()
                        t.step_y()
                res.append(t.grid().copy())
        assert gd.same_bits(res[0], res[1])


def test_step_n_matches_repeated_steps():
    cases, data = gd.tiny(np.float64)
    c = cases[-3]
    k = c["key"]
    with make(data[k + "_init"], c["dt"], c["h"], c["mat"]) as t:
        t.step_n(10)
        assert gd.same_bits(t.grid(), data[k + "_step10"])


def test_config0_plain_parity():
    """512x512, 100 steps from the centred disk: state and VTK bit-identical
    to the reference."""
    meta, samples = gd.config0()
    b = meta["ball"]
    with Task(meta["nx"], meta["ny"], meta["dt"], meta["h"], meta["mat"]) as t:
        t.set_ball(b["c"], b["r"], b["s"])
        assert gd.sha256(t.view()) == b["init_sha256"]
        for _ in range(meta["steps"]):
            t.step()
        g = t.grid()
        assert gd.same_bits(g.reshape(-1, 5)[samples["index"]], samples["values"])
        assert gd.sha256(g) == meta["state_sha256"]
        with tempfile.TemporaryDirectory() as td:
            p = os.path.join(td, "c0.vtk")
            assert t.save(p) == 0
            assert hashlib.sha256(open(p, "rb").read()).hexdigest() == meta["vtk_sha256"]


RAGGED = [(3, 3), (5, 257), (257, 5), (1000, 7), (7, 1000), (333, 129), (1031, 517)]


@pytest.mark.parametrize("nx,ny", RAGGED)
def test_random_fields_match_oracle(nx, ny):
    rng = np.random.default_rng(nx * 7919 + ny)
    g0 = rng.standard_normal((ny, nx, 5))
    g0[:, : nx // 3] = 0.5  # flat region: eps/eps ratios
    dt, h, mat = 0.13, (0.9, 1.1), (2.4, 1.3, 1.7)
    ref = g0.copy()
    OD.run(ref, dt, h, mat, 3)
    with make(g0, dt, h, mat) as t:
        t.step_n(3)
        assert gd.same_bits(t.grid(), ref)


@pytest.mark.parametrize("nx,ny", [(300, 200), (70, 513)])
def test_ring_edge_entry_points(nx, ny):
    rng = np.random.default_rng(11)
    g0 = rng.standard_normal((ny, nx, 5))
    dt, h, mat = 0.2, (1.0, 1.0), (2.0, 1.0, 1.5)
    ref = g0.copy()
    OD.sweep_x(ref, dt, h[0], mat)
    OD.sweep_y(ref, dt, h[1], mat, edge=EDGE_RING)
    with make(g0, dt, h, mat) as t:
        t.step_cf(7)
        assert gd.same_bits(t.grid(), ref)
    soa = np.ascontiguousarray(np.moveaxis(g0, -1, 0))
    with make(soa, dt, h, mat, ST_SOA) as t:
        t.step_x()
        t.step_y()
        assert gd.same_bits(np.moveaxis(t.grid(), 0, -1), ref)


def test_tile_shapes_give_identical_bits():
    """Any launch shape must give the same bits (catches halo and LDS races)."""
    rng = np.random.default_rng(12)
    nx, ny = 1100, 700
    g0 = rng.standard_normal((ny, nx, 5))
    out = []
    for shape in [(1, 7, 64, 1, 64), (4, 64, 128, 5, 128), (16, 1000, 256, 64, 256),
                  (3, 33, 256, 1000, 128), (2, 9, 64, 3, 256)]:
        with make(g0, 0.2, (1.0, 1.0), (2.0, 1.0, 1.5)) as t:
            t.set_tiles(*shape)
            assert t.get_tiles() == shape
            t.step_n(2)  # fused step kernel
            for _ in range(2):  # separate sweep kernels
                t.step_x()
                t.step_y()
            out.append(t.grid().copy())
    for o in out[1:]:
        assert gd.same_bits(o, out[0])
    ref = g0.copy()
    OD.run(ref, 0.2, (1.0, 1.0), (2.0, 1.0, 1.5), 4)
    assert gd.same_bits(out[0], ref)


def test_sync_modes_and_host_edits():
    rng = np.random.default_rng(13)
    g0 = rng.standard_normal((40, 50, 5))
    dt, h, mat = 0.2, (1.0, 1.0), (2.0, 1.0, 1.5)
    ref = g0.copy()
    OD.run(ref, dt, h, mat, 2)
    with make(g0, dt, h, mat) as t:
        t.set_sync_mode(le2d_amd.SYNC_ASYNC)
        t.step()
        t.step()
        t.wait()
        assert gd.same_bits(t.grid(), ref)
    with make(g0, dt, h, mat) as t:
        t.set_sync_mode(le2d_amd.SYNC_COHERENT)
        t.step()
        t.step()
        assert gd.same_bits(t.view(), ref)  # no explicit sync needed
    # host edit between steps, declared with le_mark_host_dirty
    ref2 = g0.copy()
    OD.run(ref2, dt, h, mat, 1)
    ref2[10, 10, :] = 3.0
    OD.run(ref2, dt, h, mat, 1)
    with make(g0, dt, h, mat) as t:
        t.step()
        g = t.grid()
        g[10, 10, :] = 3.0
        t.lib.le_mark_host_dirty(t.ref)
        t.step()
        assert gd.same_bits(t.grid(), ref2)


def test_interleaved_tasks_are_independent():
    rng = np.random.default_rng(14)
    a0 = rng.standard_normal((31, 45, 5))
    b0 = rng.standard_normal((64, 17, 5))
    ra, rb = a0.copy(), b0.copy()
    OD.run(ra, 0.2, (1.0, 1.0), (2.0, 1.0, 1.5), 3)
    OD.run(rb, 0.1, (0.5, 0.7), (3.0, 1.5, 2.0), 3)
    with make(a0, 0.2, (1.0, 1.0), (2.0, 1.0, 1.5)) as ta, \
            make(b0, 0.1, (0.5, 0.7), (3.0, 1.5, 2.0)) as tb:
        for _ in range(3):
            ta.step()
            tb.step()
        assert gd.same_bits(ta.grid(), ra)
        assert gd.same_bits(tb.grid(), rb)


def test_set_ball_after_steps_sees_current_state():
    with Task(64, 64, 0.25) as t:
        t.set_ball((32.0, 32.0), 6.0, 1.0)
        t.step_n(3)
        t.set_ball((10.0, 50.0), 4.0, -1.0)  # must not drop the 3 steps
        t.step()
        got = t.grid().copy()
    ref = np.zeros((64, 64, 5))
    OD.set_ball(ref, (1.0, 1.0), (32.0, 32.0), 6.0, 1.0)
    OD.run(ref, 0.25, (1.0, 1.0), (2.0, 1.0, 1.0), 3)
    OD.set_ball(ref, (1.0, 1.0), (10.0, 50.0), 4.0, -1.0)
    OD.run(ref, 0.25, (1.0, 1.0), (2.0, 1.0, 1.0), 1)
    assert gd.same_bits(got, ref)


def test_tasks_stepped_from_two_threads():
    """Distinct tasks stepped concurrently from two host threads (one stream
    per task; the segment tuner and its cache shared) each stay
    bit-identical to the oracle."""
    import threading
    specs = [((300, 257), 0.2, (1.0, 1.0), (2.0, 1.0, 1.5), 21),
             ((1024, 96), 0.15, (0.9, 1.1), (2.6, 1.4, 1.2), 17)]
    rng = np.random.default_rng(16)
    inits = [rng.standard_normal((ny, nx, 5)) for (nx, ny), *_ in specs]
    out, errs = [None, None], []

    def work(k):
        try:
            (nx, ny), dt, h, mat, steps = specs[k]
            with make(inits[k], dt, h, mat) as t:
                t.step_n(steps - 5)  # long enough for the segment tuner
                for _ in range(5):
                    t.step()  # calls paired into two-step passes
                out[k] = t.grid().copy()
        except Exception as e:  # reported by the main thread
            errs.append(e)
    th = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for x in th:
        x.start()
    for x in th:
        x.join(120)
    assert not errs, errs
    for k, ((nx, ny), dt, h, mat, steps) in enumerate(specs):
        ref = inits[k].copy()
        OD.run(ref, dt, h, mat, steps)
        assert gd.same_bits(out[k], ref), specs[k][0]


def test_long_trajectory_matches_oracle():
    """60 steps on a ragged 1031x517 grid from two disks over opposite
    corners: many two-step passes (tuned segment), fronts on the clamped
    edges from the start, and the zero-state limiter cases between them,
    bit-identical to the oracle."""
    nx, ny, dt, h, mat = 1031, 517, 0.2, (1.0, 0.8), (2.2, 1.2, 1.4)
    with Task(nx, ny, dt, h, mat) as t:
        t.set_ball((30.0, 30.0), 50.0, 1.0)
        t.set_ball((1000.0, 400.0), 40.0, -0.5)
        ref = t.grid().copy()
        t.step_n(60)
        got = t.grid().copy()
    OD.run(ref, dt, h, mat, 60)
    # the fronts reached every edge
    assert all(np.abs(e).max() > 0 for e in (ref[0], ref[-1], ref[:, 0], ref[:, -1]))
    assert gd.same_bits(got, ref)


@pytest.mark.parametrize("nx,ny", [(3, 70001), (70001, 3), (4, 131075)])
def test_extreme_aspect_ratios_match_oracle(nx, ny):
    """Grids of a few columns and many rows (the y-sweep's segment count
    passes the 65535 limit of a launch dimension) and of a few rows: fused
    steps, two-step passes and single sweeps against the oracle."""
    rng = np.random.default_rng(nx + ny)
    g0 = rng.standard_normal((ny, nx, 5))
    dt, h, mat = 0.2, (1.0, 0.9), (2.0, 1.0, 1.5)
    ref = g0.copy()
    OD.run(ref, dt, h, mat, 5)
    with make(g0, dt, h, mat) as t:
        t.step_n(5)
        assert gd.same_bits(t.grid(), ref)
    rx = g0.copy()
    OD.sweep_x(rx, dt, h[0], mat)
    ry = rx.copy()
    OD.sweep_y(ry, dt, h[1], mat)
    with make(g0, dt, h, mat) as t:
        t.set_tiles(y_seg=1)
        t.step_x()
        assert gd.same_bits(t.grid(), rx)
        t.step_y()
        assert gd.same_bits(t.grid(), ry)
