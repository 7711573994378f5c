"""The C-ABI libraries without a GPU: they load, export every symbol the
headers declare, keep the reference's struct layout, and their host-side entry
points (material, disk, VTK) match the reference's golden outputs.  Stepping
without a device must fail loudly (there is no CPU path)."""
import ctypes
import os
import re
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import golden_data as gd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import le2d_amd  # noqa: E402
from le2d_amd.abi import ABI_D, ABI_F  # noqa: E402

HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"static __inline__.*?\n}\n", "", text, flags=re.S)
        text = re.sub(r"^typedef [^;{]*;", "", text, flags=re.M)  # function-pointer types
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([a-z][a-z0-9_]*)\s*\([^;{]*\)\s*;", text,
                             flags=re.M):
            names.add(m.group(1))
    return names


@pytest.fixture(scope="module", params=[np.float64, np.float32], ids=["d", "f"])
def lib(request):
    return le2d_amd.load(request.param)


def test_headers_declare_the_reference_api():
    names = declared_functions()
    for n in ("le_init_material", "le_init_task", "le_free_task", "le_set_ball", "le_save_task",
              "le_step", "le_step_cf", "le_step_soa", "le_step_x", "le_step_y", "le_step_y_cf",
              "le_soa_step_x", "le_soa_step_y"):
        assert n in names


def test_python_constants_match_header():
    text = open(os.path.join(ROOT, "include", "le2d.h")).read()
    defs = {m.group(1): int(m.group(2))
            for m in re.finditer(r"^#define LE2D_([A-Z_]+) (\d+)\s*$", text, flags=re.M)}
    for name in ("ERR_ARG", "ERR_DEVICE", "ERR_NOMEM", "ERR_STATE", "ERR_BUDGET", "ERR_BUSY"):
        assert getattr(le2d_amd, name) == defs[name], name
    for name in ("SYNC_WAIT", "SYNC_ASYNC", "SYNC_COHERENT"):
        assert getattr(le2d_amd, name) == defs[name], name


def test_every_declared_symbol_is_exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib._name], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = declared_functions() - exported
    assert not missing, missing


@pytest.mark.parametrize("abi,size,offsets,mat", [
    (ABI_D, 112, [0, 8, 24, 32, 96, 104], 64),
    (ABI_F, 64, [0, 4, 12, 20, 52, 56], 32)])
def test_struct_layout(abi, size, offsets, mat):
    T = abi.le_task
    assert ctypes.sizeof(T) == size
    assert [getattr(T, f).offset for f in ("dt", "h", "n", "mat", "stype", "grid")] == offsets
    assert ctypes.sizeof(abi.le_material) == mat


def test_c_header_layout_matches(tmp_path):
    """Compile a C probe against include/le_core.h and compare offsets."""
    src = tmp_path / "probe.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "le_core.h"\n'
        'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(le_task),'
        'offsetof(le_task,dt),offsetof(le_task,h),offsetof(le_task,n),offsetof(le_task,mat),'
        'offsetof(le_task,stype),offsetof(le_task,grid),sizeof(le_material),sizeof(le_node));'
        'return 0;}\n')
    for flag, expect in (("-DUSE_DOUBLE", "112 0 8 24 32 96 104 64 40"),
                         ("-UUSE_DOUBLE", "64 0 4 12 20 52 56 32 20")):
        exe = tmp_path / "probe"
        subprocess.run(["gcc", flag, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       check=True)
        assert subprocess.run([str(exe)], capture_output=True, text=True).stdout.strip() == expect


def test_extension_structs_match_python(tmp_path):
    """le2d_request, le2d_server_stats and le2d_view: the C layout of
    include/le2d.h and the ctypes mirrors agree."""
    src = tmp_path / "probe2.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "le2d.h"\n'
        'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(le2d_request),'
        'offsetof(le2d_request,keep_result),sizeof(le2d_server_stats),'
        'offsetof(le2d_server_stats,cancelled),offsetof(le2d_server_stats,rejected),'
        'sizeof(le2d_view),offsetof(le2d_view,plane));return 0;}\n')
    exe = tmp_path / "probe2"
    subprocess.run(["gcc", "-DUSE_DOUBLE", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    R, S, V = le2d_amd.le2d_request, le2d_amd.le2d_server_stats, le2d_amd.le2d_view
    assert got == [ctypes.sizeof(R), R.keep_result.offset, ctypes.sizeof(S), S.cancelled.offset,
                   S.rejected.offset, ctypes.sizeof(V), V.plane.offset]


def test_allocator_hook_arguments(lib):
    """Both allocator functions or neither: one alone is refused; both NULL
    restores hipMalloc / hipFree (no GPU needed)."""
    fn = le2d_amd.ALLOC_FN(lambda p, n, c: 1)
    assert lib.le2d_set_device_allocator(ctypes.cast(fn, ctypes.c_void_p), None, None) == \
        le2d_amd.ERR_ARG
    assert lib.le2d_set_device_allocator(None, None, None) == 0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_material_matches_reference(dtype):
    lib = le2d_amd.load(dtype)
    m = gd.materials()
    names = ("c1", "c2", "rho", "irhoc1", "irhoc2", "rhoc1", "rhoc2", "rhoc3")
    for inp, exp in zip(m["inputs"], m[np.dtype(dtype).name]):
        mat = lib.abi.le_material()
        lib.le_init_material(*[gd.fromhex(x) for x in inp], ctypes.byref(mat))
        assert [getattr(mat, n) for n in names] == [gd.fromhex(x) for x in exp]


def test_set_ball_matches_reference():
    for c in gd.balls():
        dt = np.float64 if c["dtype"] == "float64" else np.float32
        nx, ny = c["n"]
        with le2d_amd.Task(nx, ny, 0.1, c["h"], (2.0, 1.0, 1.0), dtype=dt) as t:
            t.set_ball(c["c"], c["r"], c["s"])
            assert gd.sha256(t.view()) == c["sha256"]


def test_set_ball_soa_storage():
    with le2d_amd.Task(40, 30, 0.1, (1.0, 1.0), stype=le2d_amd.ST_SOA) as t:
        t.set_ball((20.0, 15.0), 10.0, 1.0)
        v = t.view()
        assert v.shape == (5, 30, 40)
        ref = [c for c in gd.balls() if c["dtype"] == "float64" and c["n"] == [40, 30]
               and c["h"] == [1.0, 1.0]][0]
        assert np.argwhere(v[2] != 0).tolist() == ref["nodes"]
        assert np.array_equal(v[2], v[4]) and not v[(0, 1, 3),].any()


def test_save_task_matches_reference(golden_dir, tmp_path):
    state = np.load(os.path.join(golden_dir, "tiny_vtk_d_state.npy"))
    expect = open(os.path.join(golden_dir, "tiny_vtk_d.vtk"), "rb").read()
    with le2d_amd.Task(13, 7, 0.2, (0.5, 2.25)) as t:
        t.load(state)
        p = str(tmp_path / "t.vtk")
        assert t.save(p) == 0
        assert open(p, "rb").read() == expect
        assert t.save(str(tmp_path / "no" / "such" / "dir.vtk")) == 1


def test_step_without_gpu_fails_loudly():
    """No CPU fallback: a step with no HIP device aborts with a message."""
    if le2d_amd_has_device():
        pytest.skip("a HIP device is present")
    code = ("import sys; sys.path.insert(0, %r); import le2d_amd; "
            "t = le2d_amd.Task(8, 8, 0.1); t.step()" % ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "no HIP device" in r.stderr


def le2d_amd_has_device():
    return os.path.exists("/dev/kfd")
