"""Two steps per pass (le2d_set_step_pairs): bit-identical to single steps for
every size (one and several strips, grids narrower than a window, one to a
few rows), segment height (1 row to more than the grid), window width (64,
128 and 256 lanes; 64-lane windows with the stages on one or two
wavefronts) and element type,
odd step counts included (the last step takes the single-step pass), and
to the oracle in FP64."""
import numpy as np
import pytest

import golden_data as gd
import le2d_amd
from oracle import Oracle

pytestmark = pytest.mark.gpu
SIZES = [(3, 3), (4, 5), (7, 3), (16, 9), (33, 20), (250, 7), (300, 257), (517, 129), (1031, 67),
         # widths at and one past the strip widths of 64-, 128- and 256-lane windows
         (56, 9), (57, 11), (112, 6), (120, 5), (121, 7), (248, 4), (249, 5), (496, 3)]
SEGS = (0, 1, 3, 7, 32, 1000)
ARGS = (0.2, (1.0, 0.9), (2.0, 1.0, 1.5))


def _run(g0, dtype, seg, steps, lanes=256, split=0):
    ny, nx = g0.shape[:2]
    with le2d_amd.Task(nx, ny, *ARGS, dtype=dtype) as t:
        t.load(g0)
        t.set_step_pairs(seg)
        t.set_step_pair_window(lanes)
        t.set_step_pair_split(split)
        assert t.get_step_pairs() == seg
        assert t.get_step_pair_window() == lanes
        assert t.get_step_pair_split() == split
        t.step_n(steps)
        return t.grid().copy()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nx,ny", SIZES)
def test_pairs_equal_single_steps(dtype, nx, ny):
    rng = np.random.default_rng(1000 * nx + ny)
    g0 = rng.standard_normal((ny, nx, 5)).astype(dtype)
    single = _run(g0, dtype, 0, 5)
    for lanes, split in ((256, 0), (128, 0), (64, 0), (64, 1)):
        for seg in SEGS[1:]:
            assert gd.same_bits(_run(g0, dtype, seg, 5, lanes, split), single), (lanes, split, seg)


@pytest.mark.parametrize("lanes,split", [(256, 0), (64, 0), (64, 1)])
def test_pairs_match_oracle(lanes, split):
    rng = np.random.default_rng(5)
    g0 = rng.standard_normal((131, 517, 5))
    got = _run(g0, np.float64, 16, 6, lanes, split)
    ref = g0.copy()
    Oracle(np.float64).run(ref, *ARGS, 6)
    assert gd.same_bits(got, ref)


def test_pairs_disk_workload_and_timing_entry():
    n = 1024
    with le2d_amd.Task(n, n, 0.25, (1.0, 1.0), (2.0, 1.0, 1.0)) as t:
        t.set_ball((500.0, 520.0), 64.0, 1.0)
        ref = t.grid().copy()
        t.set_step_pairs(32)
        t.step_n(10)
        got = t.grid().copy()
        assert t.time_pairs(2) > 0  # advances 4 more steps
    Oracle(np.float64).run(ref, 0.25, (1.0, 1.0), (2.0, 1.0, 1.0), 10)
    assert gd.same_bits(got, ref)


def test_tuning_leaves_the_state_alone():
    rng = np.random.default_rng(11)
    g0 = rng.standard_normal((300, 700, 5))
    with le2d_amd.Task(700, 300, *ARGS) as t:
        t.load(g0)
        t.step_n(3)
        before = t.grid().copy()
        seg = t.tune_step_pairs()
        assert seg in (4, 8, 12, 16, 24, 32, 48, 64, 80, 96, 128, 160)
        assert t.get_step_pairs() == seg
        assert t.get_step_pair_window() in (64, 128, 256)
        assert gd.same_bits(t.grid(), before)
        t.step_n(9)
        got = t.grid().copy()
    assert gd.same_bits(got, _run(g0, np.float64, 0, 12))
