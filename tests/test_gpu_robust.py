"""Failure handling of the device-resident paths: a bounded serving queue
(back-pressure), request cancellation, kept results released exactly once,
admission that lets later requests pass one that does not fit, and device
allocation failures injected through the allocator hook
(le2d_set_device_allocator).  Every failure is forced from the host; every
kernel that runs exits normally."""
import ctypes

import numpy as np
import pytest

import golden_data as gd
import le2d_amd
from oracle import Oracle

pytestmark = pytest.mark.gpu
OD = Oracle(np.float64)
MB = 1 << 20

_hip = ctypes.CDLL("libamdhip64.so")
_hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
_hip.hipMalloc.restype = ctypes.c_int
_hip.hipFree.argtypes = [ctypes.c_void_p]
_hip.hipFree.restype = ctypes.c_int


class Hook:
    """A device allocator that counts, refuses on demand and otherwise
    delegates to hipMalloc / hipFree."""

    def __init__(self):
        self.allocs, self.frees, self.refused = 0, 0, 0
        self.refuse_from = None  # refuse the n-th allocation (0-based) and later ones
        self.live = set()
        self._a = le2d_amd.ALLOC_FN(self._alloc)
        self._f = le2d_amd.FREE_FN(self._free)

    def _alloc(self, ptr, nbytes, ctx):
        if self.refuse_from is not None and self.allocs >= self.refuse_from:
            self.refused += 1
            return 1
        rc = _hip.hipMalloc(ptr, nbytes)
        if rc == 0:
            self.allocs += 1
            self.live.add(ptr[0])
        return rc

    def _free(self, p, ctx):
        self.frees += 1
        self.live.discard(p)
        _hip.hipFree(p)

    def install(self, dtype=np.float64):
        le2d_amd.set_device_allocator(self._a, self._f, dtype)


@pytest.fixture
def hook():
    h = Hook()
    yield h
    for dt in (np.float64, np.float32):
        le2d_amd.set_device_allocator(None, None, dt)


def _disk(n=96):
    g0 = np.zeros((n, n + 16, 5))
    OD.set_ball(g0, (1.0, 1.0), (40.0, 50.0), 20.0, 1.0)
    return g0


def test_allocator_hook_routes_every_allocation(hook):
    """State buffers, material tables and reduction partials come from the
    installed allocator and go back to it; the results are unchanged."""
    hook.install()
    g0 = _disk()
    ny, nx = g0.shape[:2]
    ids = np.zeros((ny, nx), np.uint8)
    ids[:, nx // 2:] = 1
    tab = [(2.0, 1.0, 1.0), (2.6, 1.4, 2.2)]
    with le2d_amd.Task(nx, ny, 0.5 / 2.6, (1.0, 1.0), (2.0, 1.0, 1.0)) as t:
        t.load(g0)
        t.step_n(4)
        t.energy()
        t.set_materials(tab, ids)
        t.step_n(3)
        got = t.grid().copy()
    ref = g0.copy()
    OD.run(ref, 0.5 / 2.6, (1.0, 1.0), (2.0, 1.0, 1.0), 4)
    OD.run_mat(ref, 0.5 / 2.6, (1.0, 1.0), tab, ids, 3)
    assert gd.same_bits(got, ref)
    assert hook.allocs >= 4 and hook.frees == hook.allocs and not hook.live


def test_refused_allocation_leaves_the_task_usable(hook):
    """A refused state allocation fails le_step_n with LE2D_ERR_NOMEM; with
    memory back, the same task steps as if nothing had happened."""
    g0 = _disk()
    ny, nx = g0.shape[:2]
    with le2d_amd.Task(nx, ny, 0.2, (1.0, 1.0), (2.0, 1.0, 1.5)) as t:
        t.load(g0)
        hook.refuse_from = 0
        hook.install()
        rc = t.lib.le_step_n(t.ref, 3)
        assert rc == le2d_amd.ERR_NOMEM and hook.refused >= 1
        assert "refused" in t.lib.le2d_last_error().decode()
        hook.refuse_from = None
        t.step_n(3)
        got = t.grid().copy()
    ref = g0.copy()
    OD.run(ref, 0.2, (1.0, 1.0), (2.0, 1.0, 1.5), 3)
    assert gd.same_bits(got, ref)
    assert hook.frees == hook.allocs


@pytest.mark.parametrize("fail_at", [0, 1, 2])
def test_refused_allocation_in_batch_commit(hook, fail_at):
    """Refusing the arena, the descriptor table or a tile table fails the
    commit with LE2D_ERR_NOMEM and frees what it took; the batch then commits
    and steps normally."""
    hook.install()
    hook.refuse_from = fail_at
    b = le2d_amd.Batch(np.float64)
    sizes = [(40, 30), (77, 51)]
    for nx, ny in sizes:
        b.add(nx, ny, 0.2, (1.0, 1.0), (2.0, 1.0, 1.5))
    with pytest.raises(le2d_amd.LE2DError) as e:
        b.commit()
    assert f": {le2d_amd.ERR_NOMEM} " in str(e.value)
    hook.refuse_from = None
    b.commit()
    with b:
        states = []
        for k, (nx, ny) in enumerate(sizes):
            g = np.random.default_rng(k).standard_normal((ny, nx, 5))
            b.load(k, g)
            states.append(g)
        b.step(3)
        for k, g in enumerate(states):
            ref = g.copy()
            OD.run(ref, 0.2, (1.0, 1.0), (2.0, 1.0, 1.5), 3)
            assert gd.same_bits(b.read(k), ref)
    assert hook.frees == hook.allocs and not hook.live


def test_refused_server_arena(hook):
    hook.refuse_from = 0
    hook.install()
    with pytest.raises(le2d_amd.LE2DError):
        le2d_amd.Server(64 * MB, 4)


def test_refused_table_fails_the_server_cleanly(hook):
    """The engine's table allocation refused mid-run: every waiter gets
    LE2D_ERR_DEVICE with the reason, and the server shuts down cleanly."""
    hook.install()
    with le2d_amd.Server(64 * MB, 4) as s:
        hook.refuse_from = hook.allocs  # the arena is taken; refuse what follows
        i = s.submit(64, 48, 10, np.float64)
        with pytest.raises(le2d_amd.LE2DError) as e:
            s.wait(i, 60)
        assert "refused" in str(e.value)
    assert hook.frees == hook.allocs and not hook.live


def test_queue_limit_gives_back_pressure():
    """With a queue limit, submit refuses (BusyError) instead of queueing
    without bound; accepted requests all complete."""
    with le2d_amd.Server(256 * MB, 1) as s:
        s.set_admission(max_queue=2)
        ids, busy = [], 0
        for _ in range(6):
            try:
                ids.append(s.submit(512, 512, 2000, np.float64))  # tens of ms each
            except le2d_amd.BusyError:
                busy += 1
        assert busy >= 1 and 2 <= len(ids) <= 3
        s.drain(120)
        st = s.stats()
        assert st["completed"] == len(ids) and st["rejected"] == busy
        s.set_admission(max_queue=0)
        s.submit(32, 32, 2, np.float64)
        s.drain(60)


def test_cancel_queued_live_and_kept():
    """A queued request leaves the queue, a live one stops at the end of its
    chunk and frees its memory, a kept result is discarded; a request
    running beside them still gives the oracle's bits."""
    with le2d_amd.Server(256 * MB, 2) as s:
        s.set_shape(256, 16, 4)
        live = s.submit(512, 512, 100000, np.float64)          # would run for minutes
        probe = s.submit(96, 80, 30, np.float64, dt=0.2, mat=(2.0, 1.0, 1.5),
                         disk=(40.0, 30.0, 12.0, 1.0), keep=True)
        queued = s.submit(64, 64, 10, np.float64)              # waits: both slots taken
        s.cancel(queued)
        with pytest.raises(le2d_amd.LE2DError):
            s.wait(queued, 10)
        s.wait(probe, 120)
        got = s.result(probe)
        with pytest.raises(le2d_amd.LE2DError):  # released: the result is gone
            s.result(probe)
        s.cancel(live)
        with pytest.raises(le2d_amd.LE2DError):
            s.wait(live, 120)
        kept = s.submit(48, 40, 4, np.float64, keep=True)
        s.wait(kept, 60)
        s.cancel(kept)
        with pytest.raises(le2d_amd.LE2DError):
            s.result(kept)
        with pytest.raises(le2d_amd.LE2DError):  # twice
            s.cancel(kept)
        s.drain(60)
        st = s.stats()
        assert st["cancelled"] == 3 and st["bytes_in_use"] == 0
    g = np.zeros((80, 96, 5))
    OD.set_ball(g, (1.0, 1.0), (40.0, 30.0), 12.0, 1.0)
    OD.run(g, 0.2, (1.0, 1.0), (2.0, 1.0, 1.5), 30)
    assert gd.same_bits(got, g)


def test_admission_passes_a_request_that_does_not_fit():
    """A request too large for the free budget waits; smaller ones behind it
    are admitted meanwhile (up to max_pass times), and it runs once memory
    frees up.  With max_pass = 0 admission is strictly in arrival order."""
    for max_pass, small_first in ((16, True), (0, False)):
        with le2d_amd.Server(128 * MB, 8) as s:
            s.set_admission(max_queue=0, window=64, max_pass=max_pass)
            s.set_shape(256, 16, 4)
            a = s.submit(1024, 1024, 40, np.float64)   # 2 x 40 MiB of fields: fits alone
            big = s.submit(1024, 1024, 4, np.float64)  # does not fit beside `a` in 128 MiB
            small = s.submit(64, 64, 4, np.float64)
            s.drain(120)
            t_big, t_small = s.times(big)[1], s.times(small)[1]
            assert (t_small < t_big) == small_first, (max_pass, t_small, t_big)
            assert s.times(a)[2] <= t_big + 1e-3


def test_grids_below_three_nodes_a_line_are_refused():
    """The sweeps need three nodes per line (the reference forms the
    invariants of nodes 0..2 before its loop, simple/le_core.c:482-486, and
    reads past the line otherwise): a task, a batch instance or a served
    request with nx or ny below 3 is refused with LE2D_ERR_ARG, and the
    object stays usable."""
    for nx, ny in ((2, 5), (5, 2), (1, 1)):
        with le2d_amd.Task(nx, ny, 0.2) as t:
            with pytest.raises(le2d_amd.LE2DError, match=r"^le_step_n: 1 .*>= 3"):
                t.step_n(2)
    with le2d_amd.Batch(np.float64) as b:
        with pytest.raises(le2d_amd.LE2DError, match="^le2d_batch_add: 1 "):
            b.add(2, 9, 0.2)
        assert b.add(3, 3, 0.2) == 0
    with le2d_amd.Server(64 * MB, 4) as s:
        with pytest.raises(le2d_amd.LE2DError, match="^le2d_server_submit: 1 "):
            s.submit(9, 2, 10)
        i = s.submit(9, 9, 4, np.float64, keep=True)
        s.wait(i)
        assert s.result(i).shape == (9, 9, 5)
