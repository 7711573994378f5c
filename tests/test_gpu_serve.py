"""Serving loop (le2d_server_*): admission by HBM budget and slots, continuous
batching; each result is bit-identical alone and in a mix of other requests,
and FP64 results are bit-identical to the oracle."""
import threading
import time

import numpy as np
import pytest

import golden_data as gd
import le2d_amd
from oracle import Oracle

pytestmark = pytest.mark.gpu
OD = Oracle(np.float64)
GB = 1 << 30

REQ = dict(nx=300, ny=257, steps=37, dt=0.12, h=(1.0, 0.9), mat=(2.7, 1.3, 1.9),
           disk=(120.5, 101.0, 20.0, 1.0))


def _others(rng, k):
    out = []
    for _ in range(k):
        nx, ny = (int(rng.integers(16, 400)) for _ in range(2))
        c1 = float(rng.uniform(1, 4))
        out.append(dict(nx=nx, ny=ny, steps=int(rng.integers(1, 60)),
                        dtype=np.float32 if rng.random() < 0.7 else np.float64,
                        dt=0.45 / c1, mat=(c1, c1 / float(rng.uniform(1.5, 2.2)),
                                           float(rng.uniform(1, 3))),
                        disk=(nx / 3, ny / 2, float(rng.uniform(2, 10)), 1.0)))
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_alone_equals_mix(dtype):
    with le2d_amd.Server(1 * GB, 64) as s:
        i = s.submit(**REQ, dtype=dtype, keep=True)
        s.wait(i, 60)
        alone = s.result(i)
    rng = np.random.default_rng(9)
    with le2d_amd.Server(1 * GB, 64) as s:
        s.set_shape(128, 16, 5)
        ids = [s.submit(**o) for o in _others(rng, 10)]
        i = s.submit(**REQ, dtype=dtype, keep=True)
        done = []

        def late():  # requests arriving while the probe is in flight
            for o in _others(rng, 12):
                done.append(s.submit(**o))
                time.sleep(0.002)
        th = threading.Thread(target=late)
        th.start()
        th.join()
        s.wait(i, 60)
        mixed = s.result(i)
        s.drain(60)
        st = s.stats()
        assert st["completed"] == len(ids) + len(done) + 1
        assert st["peak_live"] > 1
    assert gd.same_bits(alone, mixed)


def test_fp64_request_matches_oracle():
    with le2d_amd.Server(1 * GB, 8) as s:
        i = s.submit(**REQ, dtype=np.float64, keep=True)
        s.wait(i, 60)
        got = s.result(i)
    ref = np.zeros((257, 300, 5))
    OD.set_ball(ref, REQ["h"], REQ["disk"][:2], REQ["disk"][2], REQ["disk"][3])
    OD.run(ref, REQ["dt"], REQ["h"], REQ["mat"], REQ["steps"])
    assert gd.same_bits(got, ref)


def test_serial_plain_server_equals_batched():
    res = []
    for plain in (False, True):
        with le2d_amd.Server(1 * GB, 16, plain=plain) as s:
            ids = [s.submit(**REQ, dtype=dt, keep=True) for dt in (np.float32, np.float64)]
            s.drain(60)
            res.append([s.result(i) for i in ids])
            if plain:
                assert s.stats()["peak_live"] == 1
    for a, b in zip(*res):
        assert gd.same_bits(a, b)


def test_budget_limits_admission_and_rejects_oversize():
    # a 512x512 FP64 request holds two states of ~10.5 MB each
    budget = 48 << 20
    with le2d_amd.Server(budget, 64) as s:
        ids = [s.submit(nx=512, ny=512, steps=20, dtype=np.float64) for _ in range(6)]
        s.drain(60)
        st = s.stats()
        assert st["completed"] == 6
        assert st["peak_live"] == 2
        assert st["peak_bytes"] <= budget
        assert st["bytes_in_use"] == 0
        with pytest.raises(le2d_amd.LE2DError):
            s.submit(nx=4096, ny=4096, steps=1, dtype=np.float64)
        for i in ids:
            ts, ta, td = s.times(i)
            assert ts <= ta <= td


def test_two_step_launches_change_nothing():
    # odd chunk limit and odd step count: two-step launches plus single ones
    res = []
    for pairs, lanes, split in ((0, 0, 0), (64, 0, 0), (5, 64, 0), (16, 128, 0), (7, 256, 0),
                                (5, 64, 1), (64, 64, 1)):
        with le2d_amd.Server(1 * GB, 16) as s:
            s.set_pairs(pairs)
            s.set_pair_window(lanes)
            s.set_pair_split(split)
            s.set_shape(256, 16, 7)
            ids = [s.submit(**REQ, dtype=dt, keep=True) for dt in (np.float32, np.float64)]
            s.drain(60)
            res.append([s.result(i) for i in ids])
    for other in res[1:]:
        for a, b in zip(res[0], other):
            assert gd.same_bits(a, b)


@pytest.mark.parametrize("chunk", [8, 5])
def test_every_request_of_a_ragged_mix_matches_oracle(chunk):
    # requests that finish inside a chunk (odd and even step counts, one
    # step, steps beyond several chunks) each take exactly their own steps:
    # every kept result is bit-identical to the oracle
    rng = np.random.default_rng(31 + chunk)
    specs = []
    for steps in (1, 2, 3, 7, 8, 9, 13, 16, 17, 30, 41):
        nx, ny = (int(rng.integers(24, 160)) for _ in range(2))
        c1 = float(rng.uniform(1, 4))
        specs.append(dict(nx=nx, ny=ny, steps=steps, dt=0.45 / c1, h=(1.0, 1.0),
                          mat=(c1, c1 / float(rng.uniform(1.5, 2.2)), float(rng.uniform(1, 3))),
                          disk=(nx / 2.0, ny / 3.0, float(rng.uniform(3, 9)), 1.0)))
    with le2d_amd.Server(1 * GB, 64) as s:
        s.set_shape(64, 16, chunk)
        ids = [s.submit(**q, dtype=np.float64, keep=True) for q in specs]
        s.drain(60)
        got = [s.result(i) for i in ids]
    for q, g in zip(specs, got):
        ref = np.zeros((q["ny"], q["nx"], 5))
        cx, cy, r, sv = q["disk"]
        OD.set_ball(ref, q["h"], (cx, cy), r, sv)
        OD.run(ref, q["dt"], q["h"], q["mat"], q["steps"])
        assert gd.same_bits(g, ref), (q["nx"], q["ny"], q["steps"])


def test_settings_changed_while_running():
    # launch settings changed from another thread while requests run take
    # effect at a table rebuild; results stay bit-identical to the oracle
    rng = np.random.default_rng(77)
    specs = []
    for steps in (11, 24, 37, 50):
        nx, ny = (int(rng.integers(64, 200)) for _ in range(2))
        specs.append(dict(nx=nx, ny=ny, steps=steps, dt=0.2, h=(1.0, 1.0), mat=(2.0, 1.0, 1.5),
                          disk=(nx / 2.0, ny / 2.0, 6.0, 1.0)))
    with le2d_amd.Server(1 * GB, 16) as s:
        s.set_shape(64, 16, 4)
        ids = [s.submit(**q, dtype=np.float64, keep=True) for q in specs]

        def toggle():
            for k in range(40):
                s.set_pair_window((64, 128, 256, 0)[k % 4])
                s.set_pair_split(k % 3 == 1)
                s.set_pairs((8, 64, 3, 16)[k % 4])
                s.set_shape((64, 128, 256)[k % 3], 16, 4)
                time.sleep(0.0005)
        th = threading.Thread(target=toggle)
        th.start()
        th.join()
        s.drain(60)
        got = [s.result(i) for i in ids]
    for q, g in zip(specs, got):
        ref = np.zeros((q["ny"], q["nx"], 5))
        cx, cy, r, sv = q["disk"]
        OD.set_ball(ref, q["h"], (cx, cy), r, sv)
        OD.run(ref, q["dt"], q["h"], q["mat"], q["steps"])
        assert gd.same_bits(g, ref), q["steps"]


def test_drained_reports_timeout_then_completion():
    with le2d_amd.Server(1 * GB, 8) as s:
        ids = [s.submit(**REQ, dtype=np.float64, keep=True) for _ in range(3)]
        first = s.drained(0.0)  # a zero wait returns at once, done or not
        assert first in (True, False)
        assert s.drained(60.0)
        assert s.stats()["completed"] == 3
        for i in ids:
            ts, ta, td = s.times(i)
            assert ts <= ta <= td
