"""Heterogeneous media (le2d_set_material_plane), FP64 energy (le2d_energy) and
device error norms (le2d_diff_norms) on the GPU.

The reference has one material per task, so the heterogeneous scheme is
pinned by (i) its exact reduction to the reference scheme -- one material, or
several identical materials on any id pattern, give the bits of le_step --
and (ii) bit parity with the oracle's restatement of the same node-local
definition (oracle/le2d_ref.c orc_step_mat).  The CPU tests pin that
restatement independently: its reduction to the golden-pinned orc_run,
node-wise equality with the homogeneous sweep in each node's material, and
the normal-incidence reflection and transmission coefficients of P and S
pulses.  The FP32 material-plane kernels are held to the FP32 tier's error
bound against the FP64 oracle."""
import numpy as np
import pytest

import golden_data as gd
import le2d_amd
from le2d_amd import ST_AOS, ST_SOA, Task
from oracle import EDGE_RING, Oracle

pytestmark = pytest.mark.gpu
OD = Oracle(np.float64)


def _task(g0, dt, h, mat, stype=ST_AOS, dtype=np.float64):
    ny, nx = (g0.shape[0], g0.shape[1]) if stype == ST_AOS else (g0.shape[1], g0.shape[2])
    t = Task(nx, ny, dt, h, mat, stype=stype, dtype=dtype)
    t.load(g0.astype(dtype))
    return t


@pytest.mark.parametrize("stype", [ST_AOS, ST_SOA])
def test_one_material_plane_is_le_step(stype):
    rng = np.random.default_rng(21)
    g0 = rng.standard_normal((97, 203, 5))
    init = g0 if stype == ST_AOS else np.ascontiguousarray(np.moveaxis(g0, -1, 0))
    mat = (2.3, 1.2, 1.7)
    with _task(init, 0.17, (1.0, 0.8), mat, stype) as a, _task(init, 0.17, (1.0, 0.8), mat, stype) as b:
        b.set_materials([mat])
        a.step_n(4)
        b.step_n(4)
        assert gd.same_bits(a.grid(), b.grid())


@pytest.mark.parametrize("xy_block", [64, 128, 256])
def test_identical_materials_random_ids_is_le_step(xy_block):
    """Every tile is 'mixed' here, so the node-local kernel itself must give
    the bits of the shared-face kernel."""
    rng = np.random.default_rng(22)
    g0 = rng.standard_normal((70, 301, 5))
    mat = (2.0, 1.0, 1.5)
    ids = rng.integers(0, 3, size=(70, 301)).astype(np.uint8)
    with _task(g0, 0.2, (1.0, 1.0), mat) as a, _task(g0, 0.2, (1.0, 1.0), mat) as b:
        b.set_tiles(xy_block=xy_block, xy_seg=11)
        b.set_materials([mat, mat, mat], ids)
        a.step_n(3)
        b.step_n(3)
        assert gd.same_bits(a.grid(), b.grid())


CASES = [((300, 200), "midline"), ((257, 131), "blobs"), ((64, 500), "rows")]


def _ids(nx, ny, kind, rng):
    if kind == "midline":
        ids = np.zeros((ny, nx), np.uint8)
        ids[:, nx // 2:] = 1
    elif kind == "blobs":
        y, x = np.mgrid[0:ny, 0:nx]
        ids = ((((x - 80) ** 2 + (y - 60) ** 2) < 900).astype(np.uint8)
               + 2 * (((x - 200) ** 2 + (y - 90) ** 2) < 400).astype(np.uint8))
    else:
        ids = np.zeros((ny, nx), np.uint8)
        ids[ny // 3:, :] = 1
        ids[2 * ny // 3:, :] = 2
    return ids


@pytest.mark.parametrize("shape,kind", CASES)
def test_heterogeneous_matches_oracle(shape, kind):
    nx, ny = shape
    rng = np.random.default_rng(nx + ny)
    tab = [(2.0, 1.0, 1.0), (3.1, 1.6, 2.4), (1.4, 0.8, 1.2)]
    ids = _ids(nx, ny, kind, rng)
    g0 = rng.standard_normal((ny, nx, 5))
    dt, h = 0.45 / 3.1, (1.0, 1.0)
    ref = g0.copy()
    OD.run_mat(ref, dt, h, tab, ids, 3)
    for xy_block, xy_seg in [(256, 17), (64, 5)]:
        with _task(g0, dt, h, (2.0, 1.0, 1.0)) as t:
            t.set_tiles(xy_block=xy_block, xy_seg=xy_seg)
            t.set_materials(tab, ids)
            t.step_n(3)
            assert gd.same_bits(t.grid(), ref), (kind, xy_block)


def test_heterogeneous_ring_edge_soa():
    nx, ny = 120, 90
    rng = np.random.default_rng(5)
    tab = [(2.0, 1.0, 1.0), (3.0, 1.5, 2.0)]
    ids = _ids(nx, ny, "midline", rng)
    g0 = rng.standard_normal((ny, nx, 5))
    ref = np.ascontiguousarray(np.moveaxis(g0, -1, 0))
    OD.run_mat(ref, 0.12, (1.0, 1.0), tab, ids, 2, edge=EDGE_RING, soa=True)
    with _task(np.ascontiguousarray(np.moveaxis(g0, -1, 0)), 0.12, (1.0, 1.0), (2.0, 1.0, 1.0),
               ST_SOA) as t:
        t.set_materials(tab, ids)
        t.step_n(2)
        assert gd.same_bits(t.grid(), ref)


def test_energy_matches_host_sum():
    rng = np.random.default_rng(6)
    g0 = rng.standard_normal((150, 230, 5))
    mat = (2.5, 1.1, 1.9)
    with _task(g0, 0.1, (0.7, 1.3), mat) as t:
        e = t.energy()
        assert e == pytest.approx(OD.energy(g0, (0.7, 1.3), mat), rel=1e-12)
        t.step_n(2)
        assert t.energy() == pytest.approx(OD.energy(t.grid(), (0.7, 1.3), mat), rel=1e-12)


def test_energy_heterogeneous_follows_the_oracle():
    """The superbee scheme does not conserve the elastic energy, nor only
    dissipate it: from a disk, the energy falls while the limiter clips the
    pulse (about -6.7% by step 60), rises by about 1.7% while it steepens
    the spreading fronts (to step 230), and falls again as the fronts leave
    through the clamped edges.  The GPU's energy of its state follows the
    oracle's energy of the oracle's state along that whole curve."""
    nx = ny = 256
    tab = [(2.0, 1.0, 1.0), (2.6, 1.4, 2.2)]
    ids = np.zeros((ny, nx), np.uint8)
    ids[:, nx // 2:] = 1
    with Task(nx, ny, 0.5 / 2.6, (1.0, 1.0), (2.0, 1.0, 1.0)) as t:
        t.set_ball((128.0, 128.0), 12.0, 1.0)
        t.set_materials(tab, ids)
        g = t.grid()
        # host energy with per-node material
        v = g.reshape(-1, 5)
        m = ids.reshape(-1)
        rho = np.array([r for _, _, r in tab])[m]
        mu = rho * np.array([c2 for _, c2, _ in tab])[m] ** 2
        lam = rho * np.array([c1 for c1, _, _ in tab])[m] ** 2 - 2 * mu
        tr, dv = v[:, 2] + v[:, 4], v[:, 2] - v[:, 4]
        e_host = float(np.sum(0.5 * rho * (v[:, 0] ** 2 + v[:, 1] ** 2) + tr * tr / (8 * (lam + mu))
                              + (0.5 * dv * dv + 2 * v[:, 3] ** 2) / (4 * mu)))
        e0 = t.energy()
        assert e0 == pytest.approx(e_host, rel=1e-12)
        ref = g.copy()
        es, eo = [e0], [OD.energy_mat(ref, (1.0, 1.0), tab, ids)]
        for _ in range(30):
            t.step_n(10)
            OD.run_mat(ref, 0.5 / 2.6, (1.0, 1.0), tab, ids, 10)
            es.append(t.energy())
            eo.append(OD.energy_mat(ref, (1.0, 1.0), tab, ids))
        assert es == pytest.approx(eo, rel=1e-12)
        r = np.array(es) / e0
        k_min = int(np.argmin(r[:20]))
        assert 4 <= k_min <= 8 and r[k_min] < 0.94, r   # falls first
        assert r[23] > r[k_min] + 0.01, r               # then rises
        assert r[-1] < r[23] - 0.05, r                  # then leaves


def test_diff_norms_across_libraries():
    rng = np.random.default_rng(7)
    g0 = rng.standard_normal((64, 96, 5))
    with _task(g0, 0.2, (1.0, 1.0), (2.0, 1.0, 1.5)) as a, \
            _task(g0, 0.2, (1.0, 1.0), (2.0, 1.0, 1.5), dtype=np.float32) as b:
        a.step_n(5)
        b.step_n(5)
        ga, gb = a.grid().copy(), b.grid().astype(np.float64)
        n = le2d_amd.diff_norms(b.device_view(), a.device_view())
        d = gb - ga
        assert n[0] == pytest.approx(np.sqrt(np.sum(d * d)), rel=1e-10)
        assert n[1] == pytest.approx(np.sqrt(np.sum(ga * ga)), rel=1e-12)
        assert n[2] == pytest.approx(np.abs(d).max(), rel=1e-12)
        assert n[3] == pytest.approx(np.abs(ga).max(), rel=1e-12)
        assert n[0] / n[1] < 1e-5


@pytest.mark.parametrize("shape,kind", CASES + [((517, 67), "checker"), ((9, 7), "blobs")])
def test_heterogeneous_pairs_equal_single_steps(shape, kind):
    # two steps per pass (single-material and node-local mixed tiles, windows
    # of 256, 128 and 64 lanes) against one step per pass and the oracle, odd
    # step count
    nx, ny = shape
    rng = np.random.default_rng(7 * nx + ny)
    tab = [(2.0, 1.0, 1.0), (3.1, 1.6, 2.4), (1.4, 0.8, 1.2)]
    if kind == "checker":
        ids = (np.add.outer(np.arange(ny) // 5, np.arange(nx) // 7) % 3).astype(np.uint8)
    else:
        ids = _ids(nx, ny, kind, rng)
    g0 = rng.standard_normal((ny, nx, 5))
    dt, h = 0.45 / 3.1, (1.0, 0.9)
    res = []
    for seg, lanes, split in ((0, 256, 0), (1, 256, 0), (6, 128, 0), (64, 256, 0), (3, 64, 0),
                              (64, 64, 0), (3, 64, 1), (64, 64, 1)):
        with _task(g0, dt, h, (2.0, 1.0, 1.0)) as t:
            t.set_materials(tab, ids)
            t.set_step_pairs(seg)
            t.set_step_pair_window(lanes)
            t.set_step_pair_split(split)
            t.step_n(5)
            res.append(t.grid().copy())
    for r in res[1:]:
        assert gd.same_bits(r, res[0])
    ref = g0.copy()
    OD.run_mat(ref, dt, h, tab, ids, 5)
    assert gd.same_bits(res[0], ref)


def test_heterogeneous_long_trajectory_matches_oracle():
    """40 steps, three materials (a midline split and a blob), waves from two
    disks crossing both interfaces; default launch shapes (two steps per
    pass), bit-identical to the oracle."""
    nx, ny = 517, 331
    y, x = np.mgrid[0:ny, 0:nx]
    ids = np.zeros((ny, nx), np.uint8)
    ids[:, nx // 2:] = 1
    ids[((x - 150) ** 2 + (y - 200) ** 2) < 60 ** 2] = 2
    tab = [(2.0, 1.0, 1.0), (3.1, 1.6, 2.4), (1.4, 0.8, 1.2)]
    dt, h = 0.45 / 3.1, (1.0, 1.0)
    g0 = np.zeros((ny, nx, 5))
    OD.set_ball(g0, h, (240.0, 170.0), 30.0, 1.0)
    OD.set_ball(g0, h, (150.0, 200.0), 12.0, -0.5)
    ref = g0.copy()
    OD.run_mat(ref, dt, h, tab, ids, 40)
    # the waves reached both interfaces
    assert np.abs(ref[:, nx // 2]).max() > 1e-3 and np.abs(ref[200, 150 + 60]).max() > 1e-3
    with _task(g0, dt, h, (2.0, 1.0, 1.0)) as t:
        t.set_materials(tab, ids)
        t.step_n(40)
        assert gd.same_bits(t.grid(), ref)


# FP32 tier bound against the FP64 oracle (relative L2 and Linf over all five
# fields); the reference's own float build stays within 6e-7 / 4e-6 of FP64
# on these cases.
FP32_L2, FP32_LINF = 1e-5, 1e-4
TAB3 = [(2.0, 1.0, 1.0), (3.1, 1.6, 2.4), (1.4, 0.8, 1.2)]


def _fp32_close(got, ref, what):
    d = got.astype(np.float64) - ref
    l2 = np.linalg.norm(d) / np.linalg.norm(ref)
    linf = np.abs(d).max() / np.abs(ref).max()
    assert l2 < FP32_L2 and linf < FP32_LINF, (what, l2, linf)


@pytest.mark.parametrize("shape,kind", CASES + [((517, 67), "checker"), ((9, 7), "blobs")])
@pytest.mark.parametrize("steps", [5, 20])
def test_heterogeneous_fp32_within_tier_bound(shape, kind, steps):
    """The FP32 material-plane kernels (single-material and node-local mixed
    tiles, one step per pass with 256- and 64-lane windows, two steps per
    pass with 256-, 128- and 64-lane windows, odd and even step counts)
    against the FP64 oracle of the node-local scheme, and the two-step
    passes against the single steps."""
    nx, ny = shape
    rng = np.random.default_rng(nx + ny)
    ids = _ids(nx, ny, kind, rng) if kind != "checker" else \
        (np.add.outer(np.arange(ny) // 5, np.arange(nx) // 7) % 3).astype(np.uint8)
    g0 = rng.standard_normal((ny, nx, 5))
    dt, h = 0.45 / 3.1, (1.0, 0.9)
    ref = g0.copy()
    OD.run_mat(ref, dt, h, TAB3, ids, steps)
    single = None
    for seg, lanes, xy_block, split in ((0, 256, 256, 0), (0, 256, 64, 0), (64, 256, 256, 0),
                                        (6, 128, 256, 0), (3, 64, 256, 0), (64, 64, 256, 0),
                                        (5, 64, 256, 1)):
        with _task(g0, dt, h, (2.0, 1.0, 1.0), dtype=np.float32) as t:
            t.set_tiles(xy_block=xy_block, xy_seg=11)
            t.set_materials(TAB3, ids)
            t.set_step_pairs(seg)
            t.set_step_pair_window(lanes)
            t.set_step_pair_split(split)
            t.step_n(steps)
            got = t.grid().copy()
        _fp32_close(got, ref, (seg, lanes, xy_block, split))
        if single is None:
            single = got
        else:
            _fp32_close(got, single.astype(np.float64), ("vs single", seg, lanes, split))


def test_heterogeneous_fp32_trajectory_across_interfaces():
    """60 FP32 steps (default launch shapes: two steps per pass) of waves from
    two disks crossing a midline and a blob interface, within the FP32 tier's
    bound of the FP64 oracle."""
    nx, ny = 517, 331
    y, x = np.mgrid[0:ny, 0:nx]
    ids = np.zeros((ny, nx), np.uint8)
    ids[:, nx // 2:] = 1
    ids[((x - 150) ** 2 + (y - 200) ** 2) < 60 ** 2] = 2
    dt, h = 0.45 / 3.1, (1.0, 1.0)
    g0 = np.zeros((ny, nx, 5))
    OD.set_ball(g0, h, (240.0, 170.0), 30.0, 1.0)
    OD.set_ball(g0, h, (150.0, 200.0), 12.0, -0.5)
    ref = g0.copy()
    OD.run_mat(ref, dt, h, TAB3, ids, 60)
    assert np.abs(ref[:, nx // 2 + 10]).max() > 1e-3  # the wave crossed the midline
    with _task(g0, dt, h, (2.0, 1.0, 1.0), dtype=np.float32) as t:
        t.set_materials(TAB3, ids)
        t.step_n(60)
        _fp32_close(t.grid().copy(), ref, "trajectory")
        e32 = t.energy()
    assert e32 == pytest.approx(OD.energy_mat(ref, h, TAB3, ids), rel=1e-5)
