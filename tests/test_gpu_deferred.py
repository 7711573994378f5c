"""le_step called in a loop: an odd step is held back and runs with the next
le_step as one two-step pass, or alone before any other use of the state.
Every sequence here must give the oracle's bits: odd and even counts, le_step
mixed with le_step_n and single sweeps, a time step changed while a step is
held, host edits and reads, the host-coherent mode, and device-side reads."""
import numpy as np
import pytest

import le2d_amd
from oracle import Oracle

pytestmark = pytest.mark.gpu
O = Oracle(np.float64)
ARGS = (0.2, (1.0, 0.9), (2.0, 1.0, 1.5))
NX, NY = 300, 140


def _g0(seed=3):
    return np.random.default_rng(seed).standard_normal((NY, NX, 5))


def _task(g0):
    t = le2d_amd.Task(NX, NY, *ARGS)
    t.load(g0)
    return t


def _bits(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.mark.parametrize("n", [1, 2, 3, 4, 7])
def test_le_step_loop(n):
    g0 = _g0()
    with _task(g0) as t:
        for _ in range(n):
            t.step()
        got = t.grid().copy()
    ref = g0.copy()
    O.run(ref, *ARGS, n)
    assert _bits(got, ref)


def test_mixed_with_step_n_and_sweeps():
    g0 = _g0(4)
    dt, h, mat = ARGS
    with _task(g0) as t:
        t.step()          # held
        t.step_n(3)       # joins: 4 steps
        t.step()          # held
        t.step_x()        # runs the held step, then one x-sweep
        t.step()
        t.step()          # a pair
        t.step()          # held, run by the read
        got = t.grid().copy()
    ref = g0.copy()
    O.run(ref, dt, h, mat, 5)
    O.sweep_x(ref, dt, h[0], mat)
    O.run(ref, dt, h, mat, 3)
    assert _bits(got, ref)


def test_time_step_changed_while_held():
    g0 = _g0(5)
    dt, h, mat = ARGS
    with _task(g0) as t:
        t.step()              # held with dt
        t.t.dt = 0.15         # the reference reads dt at every call
        t.step()
        t.step()
        got = t.grid().copy()
    ref = g0.copy()
    O.run(ref, dt, h, mat, 1)
    O.run(ref, 0.15, h, mat, 2)
    assert _bits(got, ref)


def test_host_edit_after_held_step():
    g0 = _g0(6)
    with _task(g0) as t:
        t.step()
        state = t.grid().copy()   # runs the held step
        state[10:20, 30:40, 2] += 0.5
        t.load(state)             # host edit replaces the device state
        t.step()
        t.step()
        t.step()
        got = t.grid().copy()
    ref = g0.copy()
    O.run(ref, *ARGS, 1)
    ref[10:20, 30:40, 2] += 0.5
    O.run(ref, *ARGS, 3)
    assert _bits(got, ref)


def test_coherent_mode_and_device_reads():
    g0 = _g0(7)
    with _task(g0) as t, _task(g0) as u:
        t.step()
        u.step_n(1)
        assert t.energy() == u.energy()  # the device read runs the held step
        t.step()
        t.set_sync_mode(le2d_amd.SYNC_COHERENT)  # runs it and updates the mirror
        ref = g0.copy()
        O.run(ref, *ARGS, 2)
        assert _bits(t.view(), ref)
        t.step()                                 # no deferral: mirror current
        O.run(ref, *ARGS, 1)
        assert _bits(t.view(), ref)


def test_pairing_off_by_environment(tmp_path):
    # LE2D_PAIR_CALLS=0 (read once per process): every le_step runs its own
    # step; checked in a child process against the oracle
    import os
    import subprocess
    import sys
    g0 = _g0(8)
    src, out = tmp_path / "g0.npy", tmp_path / "got.npy"
    np.save(src, g0)
    code = (
        "import sys, numpy as np; sys.path.insert(0, %r); import le2d_amd\n"
        "g0 = np.load(%r)\n"
        "with le2d_amd.Task(%d, %d, %r, %r, %r) as t:\n"
        "    t.load(g0)\n"
        "    for _ in range(3): t.step()\n"
        "    np.save(%r, t.grid())\n"
    ) % (os.path.dirname(os.path.dirname(os.path.abspath(__file__))), str(src), NX, NY,
         ARGS[0], ARGS[1], ARGS[2], str(out))
    env = dict(os.environ, LE2D_PAIR_CALLS="0")
    subprocess.run([sys.executable, "-c", code], env=env, check=True, timeout=120)
    ref = g0.copy()
    O.run(ref, *ARGS, 3)
    assert _bits(np.load(out), ref)
