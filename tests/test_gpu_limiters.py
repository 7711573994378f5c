"""The minmod and CIR builds of the HIP library (make limiters: lib_minmod/,
lib_cir/) against the oracle built with the same limiter (simple/le_core.c:
36-43 chooses it at compile time): FP64 bit for bit through every step path
(single fused steps, two-step passes on each window form, single sweeps, the
SOA ring edge, batches, material planes and served requests), FP32 within the
tier's bound of the FP64 oracle."""
import numpy as np
import pytest

import golden_data as gd
import le2d_amd
from oracle import EDGE_RING, Oracle

pytestmark = pytest.mark.gpu
LIMITERS = ["minmod", "cir"]
ARGS = (0.2, (1.0, 0.9), (2.0, 1.0, 1.5))


def _field(seed, nx, ny, dtype=np.float64):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((ny, nx, 5)).astype(dtype)


@pytest.mark.parametrize("limiter", LIMITERS)
def test_library_reports_its_limiter(limiter):
    for dt in (np.float64, np.float32):
        lib = le2d_amd.load(dt, limiter)
        assert lib.le2d_limiter() == le2d_amd.LIMITERS[limiter]


@pytest.mark.parametrize("limiter", LIMITERS)
@pytest.mark.parametrize("nx,ny", [(61, 37), (300, 131), (7, 5)])
def test_fused_steps_match_oracle(limiter, nx, ny):
    g0 = _field(31 * nx + ny, nx, ny)
    with le2d_amd.Task(nx, ny, *ARGS, limiter=limiter) as t:
        t.load(g0)
        for _ in range(3):
            t.step()
        got = t.grid().copy()
    ref = g0.copy()
    Oracle(np.float64, limiter).run(ref, *ARGS, 3)
    assert gd.same_bits(got, ref)


@pytest.mark.parametrize("limiter", LIMITERS)
@pytest.mark.parametrize("lanes,split", [(256, 0), (128, 0), (64, 0), (64, 1)])
def test_two_step_passes_match_oracle(limiter, lanes, split):
    g0 = _field(7, 517, 131)
    with le2d_amd.Task(517, 131, *ARGS, limiter=limiter) as t:
        t.load(g0)
        t.set_step_pairs(16)
        t.set_step_pair_window(lanes)
        t.set_step_pair_split(split)
        t.step_n(7)
        got = t.grid().copy()
    ref = g0.copy()
    Oracle(np.float64, limiter).run(ref, *ARGS, 7)
    assert gd.same_bits(got, ref)


@pytest.mark.parametrize("limiter", LIMITERS)
def test_single_sweeps_and_soa_match_oracle(limiter):
    o = Oracle(np.float64, limiter)
    g0 = _field(8, 97, 45)
    dt, h, mat = ARGS
    with le2d_amd.Task(97, 45, *ARGS, limiter=limiter) as t:
        t.load(g0)
        t.step_x()
        gx = t.grid().copy()
        t.step_y()
        gy = t.grid().copy()
    rx = g0.copy()
    o.sweep_x(rx, dt, h[0], mat)
    assert gd.same_bits(gx, rx)
    o.sweep_y(rx, dt, h[1], mat)
    assert gd.same_bits(gy, rx)
    p0 = np.ascontiguousarray(np.moveaxis(g0, -1, 0))
    with le2d_amd.Task(97, 45, *ARGS, stype=le2d_amd.ST_SOA, limiter=limiter) as t:
        t.load(p0)
        for _ in range(4):
            t.step()
        got = t.grid().copy()
    ref = p0.copy()
    for _ in range(4):
        o.sweep_x(ref, dt, h[0], mat, soa=True)
        o.sweep_y(ref, dt, h[1], mat, edge=EDGE_RING, soa=True)
    assert gd.same_bits(got, ref)


@pytest.mark.parametrize("limiter", LIMITERS)
def test_fp32_build_within_tier_bound(limiter):
    g0 = _field(9, 256, 200)
    with le2d_amd.Task(256, 200, *ARGS, dtype=np.float32, limiter=limiter) as t:
        t.load(g0.astype(np.float32))
        t.step_n(20)
        got = t.grid().astype(np.float64)
    ref = g0.copy()
    Oracle(np.float64, limiter).run(ref, *ARGS, 20)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert rel < 1e-5, rel


@pytest.mark.parametrize("limiter", LIMITERS)
def test_batch_matches_oracle(limiter):
    """The batched kernels of the limiter's library, both two-step forms and
    the single-step launch, against the oracle per instance."""
    sizes = [(3, 3), (64, 64), (257, 129), (130, 5), (40, 300)]
    o = Oracle(np.float64, limiter)
    for shape in [(64, 16, "pairs_split"), (128, 9, "pairs"), (256, 32, "tiles")]:
        with le2d_amd.Batch(np.float64, limiter=limiter) as b:
            states = []
            for k, (nx, ny) in enumerate(sizes):
                b.add(nx, ny, *ARGS)
            b.commit()
            for k, (nx, ny) in enumerate(sizes):
                g = _field(40 + k, nx, ny)
                b.load(k, g)
                states.append(g)
            b.set_shape(*shape)
            b.step(5)
            for k, g in enumerate(states):
                ref = g.copy()
                o.run(ref, *ARGS, 5)
                assert gd.same_bits(b.read(k), ref), (shape, sizes[k])


@pytest.mark.parametrize("limiter", LIMITERS)
def test_material_plane_matches_oracle(limiter):
    """Single-material and interface tiles of a two-material plane, one- and
    two-wavefront two-step passes, against the oracle's heterogeneous scheme
    built with the same limiter."""
    nx, ny = 150, 97
    tab = [(2.0, 1.0, 1.0), (3.1, 1.6, 2.4)]
    ids = np.zeros((ny, nx), dtype=np.uint8)
    ids[:, nx // 2:] = 1
    g0 = _field(50, nx, ny)
    dt, h = 0.45 / 3.1, (1.0, 1.0)
    ref = g0.copy()
    Oracle(np.float64, limiter).run_mat(ref, dt, h, tab, ids, 6)
    for split in (0, 1):
        with le2d_amd.Task(nx, ny, dt, h, (2.0, 1.0, 1.0), limiter=limiter) as t:
            t.load(g0)
            t.set_step_pair_window(64)
            t.set_step_pair_split(split)
            t.set_materials(tab, ids)
            t.step_n(6)
            assert gd.same_bits(t.grid(), ref), split


@pytest.mark.parametrize("limiter", LIMITERS)
def test_served_requests_match_oracle(limiter):
    """Requests of the serving loop in the limiter's library against the
    oracle: a disk of radius 10 on a zero field, as the request defines."""
    o = Oracle(np.float64, limiter)
    with le2d_amd.Server(256 << 20, 8, limiter=limiter) as s:
        ids = [s.submit(nx, ny, steps, np.float64, dt=0.2, keep=True)
               for nx, ny, steps in [(70, 50, 9), (33, 120, 6)]]
        for i, (nx, ny, steps) in zip(ids, [(70, 50, 9), (33, 120, 6)]):
            s.wait(i)
            got = s.result(i)
            ref = np.zeros((ny, nx, 5))
            o.set_ball(ref, (1.0, 1.0), (nx / 2, ny / 2), 10.0, 1.0)
            o.run(ref, 0.2, (1.0, 1.0), (2.0, 1.0, 1.0), steps)
            assert gd.same_bits(got, ref), (nx, ny, steps)
