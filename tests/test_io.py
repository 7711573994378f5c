"""le_save_task of a large grid (CPU: byte for byte the oracle writer's file),
raw dump/restore (CPU: the host mirror round-trips bit for bit) and the
command-line drivers (GPU: the reference's own simple/main.c linked to this
library, and le2d_main.c, reproduce the reference's result.vtk)."""
import hashlib
import os
import struct
import subprocess

import numpy as np
import pytest

import golden_data as gd
import le2d_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("dtype,stype", [(np.float64, le2d_amd.ST_AOS),
                                         (np.float32, le2d_amd.ST_SOA)])
def test_save_task_of_a_large_grid_is_the_oracle_writers_file(tmp_path, dtype, stype):
    """4100 x 4200 nodes: le_save_task converts the |v| rows on several
    threads, in two chunks of about 64 MB; the file is the oracle writer's
    byte for byte (CPU: a task never stepped saves its host mirror)."""
    from oracle import Oracle
    nx, ny, h = 4100, 4200, (0.7, 1.3)
    rng = np.random.default_rng(3)
    with le2d_amd.Task(nx, ny, 0.2, h, dtype=dtype, stype=stype) as t:
        v = t.view()
        if stype == le2d_amd.ST_AOS:
            v[..., :2] = rng.standard_normal((ny, nx, 2))
        else:
            v[:2] = rng.standard_normal((2, ny, nx))
        assert t.save(str(tmp_path / "a.vtk")) == 0
        g = np.ascontiguousarray(v.copy())
    assert Oracle(dtype).save_vtk(g, h, str(tmp_path / "b.vtk"),
                                  soa=stype == le2d_amd.ST_SOA) == 0
    a, b = (tmp_path / "a.vtk").read_bytes(), (tmp_path / "b.vtk").read_bytes()
    assert len(a) > (64 << 20) and a == b


@pytest.mark.parametrize("dtype,stype", [(np.float64, le2d_amd.ST_AOS),
                                         (np.float32, le2d_amd.ST_AOS),
                                         (np.float64, le2d_amd.ST_SOA)])
def test_dump_restore_roundtrip_host(tmp_path, dtype, stype):
    rng = np.random.default_rng(3)
    with le2d_amd.Task(17, 9, 0.21, (0.5, 0.75), (2.5, 1.25, 1.75), stype=stype,
                       dtype=dtype) as a:
        state = rng.standard_normal(a.view().shape).astype(dtype)
        a.load(state)
        p = str(tmp_path / "s.raw")
        a.dump(p)
        raw = open(p, "rb").read()
        assert raw[:8] == b"LE2DRAW1"
        assert struct.unpack("<4i", raw[8:24]) == (np.dtype(dtype).itemsize, 17, 9, stype)
        assert len(raw) == 128 + 5 * 17 * 9 * np.dtype(dtype).itemsize
        planes = np.frombuffer(raw[128:], dtype=dtype).reshape(5, 9, 17)
        expect = state if stype == le2d_amd.ST_SOA else np.moveaxis(state, -1, 0)
        assert np.array_equal(planes, expect)
        with le2d_amd.Task(17, 9, 1.0, (1.0, 1.0), (2.0, 1.0, 1.0), stype=stype,
                           dtype=dtype) as b:
            b.restore(p)
            assert np.array_equal(b.view(), state)
            assert (b.t.dt, b.t.h.x, b.t.h.y) == (a.t.dt, a.t.h.x, a.t.h.y)
            assert b.t.mat.rhoc3 == a.t.mat.rhoc3
        with le2d_amd.Task(16, 9, 1.0, dtype=dtype) as c:
            with pytest.raises(le2d_amd.LE2DError):
                c.restore(p)


def test_restore_rejects_bad_files_and_keeps_the_task(tmp_path):
    rng = np.random.default_rng(4)
    with le2d_amd.Task(17, 9, 0.2, dtype=np.float32) as a:
        a.load(rng.standard_normal(a.view().shape).astype(np.float32))
        good = str(tmp_path / "f.raw")
        a.dump(good)
    raw = open(good, "rb").read()
    bad = {"magic": b"XXXXRAW1" + raw[8:], "truncated": raw[:-4], "header": raw[:100]}
    for name, data in bad.items():
        with open(tmp_path / name, "wb") as f:
            f.write(data)
    with le2d_amd.Task(17, 9, 0.3, dtype=np.float32) as b:
        keep = rng.standard_normal(b.view().shape).astype(np.float32)
        b.load(keep)
        for name in bad:
            with pytest.raises(le2d_amd.LE2DError):
                b.restore(str(tmp_path / name))
        with pytest.raises(le2d_amd.LE2DError):
            b.restore(str(tmp_path / "missing.raw"))
        assert np.array_equal(b.view(), keep) and b.t.dt == np.float32(0.3)
    with le2d_amd.Task(17, 9, 0.3, dtype=np.float64) as c:  # FP32 file, FP64 task
        with pytest.raises(le2d_amd.LE2DError):
            c.restore(good)


def _run(exe, args, cwd, env=None):
    e = dict(os.environ)
    e.update(env or {})
    r = subprocess.run([exe] + args, cwd=cwd, capture_output=True, text=True, timeout=300, env=e)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _vtk_floats(path):
    data = open(path, "rb").read()
    head = data.index(b"LOOKUP_TABLE v_table\n") + len(b"LOOKUP_TABLE v_table\n")
    return data[:head], np.frombuffer(data[head:], dtype=">f4")


DROPIN = os.path.join(ROOT, "oracle", "_ref", "simple_main_on_le2d_d")
CLI_D = os.path.join(ROOT, "le2d_amd", "bin", "le2d_d")
CLI_F = os.path.join(ROOT, "le2d_amd", "bin", "le2d_f")


@pytest.mark.gpu
@pytest.mark.parametrize("args", [["64", "48", "20", "1"], ["100", "37", "7", "2"]])
def test_cli_fp64_reproduces_reference_result(tmp_path, args):
    golden = gd.cli()["d:" + " ".join(args)]
    exes = [CLI_D] + ([DROPIN] if os.path.exists(DROPIN) else [])
    for exe in exes:
        d = tmp_path / os.path.basename(exe)
        d.mkdir()
        out = _run(exe, args, str(d))
        assert out.split()[:3] == golden["stdout_prefix"].split()
        vtk = open(d / "result.vtk", "rb").read()
        assert len(vtk) == golden["vtk_bytes"]
        assert hashlib.sha256(vtk).hexdigest() == golden["vtk_sha256"], exe


@pytest.mark.gpu
def test_cli_fp32_close_to_reference_float_build(tmp_path):
    """The FP32 tier computes in float (the reference's float build promotes
    parts of tvd2 to double), so |v| agrees to a tolerance, not bit for bit."""
    args = ["64", "48", "20", "1"]
    exes = [CLI_F] + ([DROPIN[:-1] + "f"] if os.path.exists(DROPIN[:-1] + "f") else [])
    ref_dir = tmp_path / "ref"
    ref_dir.mkdir()
    # the float reference's VTK is pinned only by hash; compare to the FP64 run
    _run(CLI_D, args, str(ref_dir))
    h0, v0 = _vtk_floats(ref_dir / "result.vtk")
    for exe in exes:
        d = tmp_path / os.path.basename(exe)
        d.mkdir()
        _run(exe, args, str(d))
        h1, v1 = _vtk_floats(d / "result.vtk")
        assert h1 == h0
        np.testing.assert_allclose(v1, v0, rtol=0, atol=2e-5 * float(np.abs(v0).max()))


@pytest.mark.gpu
def test_cli_save_every_and_restart(tmp_path):
    straight = tmp_path / "straight"
    straight.mkdir()
    _run(CLI_D, ["40", "30", "12", "1"], str(straight), {"LE2D_SAVE_EVERY": "1"})
    frames = sorted(p for p in os.listdir(straight) if p.startswith("out-"))
    assert frames == [f"out-{k:06d}.vtk" for k in range(12)]
    first = tmp_path / "first"
    first.mkdir()
    dump = str(tmp_path / "mid.raw")
    _run(CLI_D, ["40", "30", "5", "1"], str(first), {"LE2D_DUMP": dump})
    second = tmp_path / "second"
    second.mkdir()
    _run(CLI_D, ["40", "30", "7", "1"], str(second), {"LE2D_RESTART": dump})
    assert open(second / "result.vtk", "rb").read() == open(straight / "result.vtk", "rb").read()
