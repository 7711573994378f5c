"""Write the golden vectors for the split-step sweeps.

Run in a container that holds the reference sources:
    make -C oracle ref && python tests/golden/make_golden.py

Every expected output here is produced by the reference's own simple/ code,
compiled from its sources (oracle/_ref/libsimple_{d,f}.so, the CLI binaries
oracle/_ref/simple_{d,f}) and called through its le_core.h entry points.
The fixtures are data only: seeded inputs and the reference's outputs.

Files written next to this script:
  tiny_d.npz / tiny_f.npz   small grids, seeded fields, states after single
                            sweeps, 1/2/10 steps, the SOA step and le_step_y_cf
  config0.json (+ config0_samples_d.npz)
                            512x512 FP64, 100 steps from the centred disk:
                            per-field SHA-256, sums, norms, 1024 sampled nodes,
                            the disk's node count and the VTK file's SHA-256
  materials.json            le_init_material outputs for seeded (c1, c2, rho)
  ball.json                 le_set_ball node sets on non-unit spacing
  cli.json                  SHA-256 of result.vtk written by the reference CLI
  tiny_vtk_d.vtk            a VTK file written by le_save_task (FP64 build)
"""
import ctypes
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import Reference  # noqa: E402

TINY_SHAPES = [(3, 3), (4, 6), (7, 5), (16, 9), (33, 20), (70, 3), (3, 41)]
# (dt, (hx, hy), (c1, c2, rho)): unit and non-unit spacing, hx != hy, Courant
# numbers from small to 1 (k1 = dt*c1/h).
TINY_CASES = [
    (0.3, (1.0, 1.0), (2.0, 1.0, 1.5)),
    (0.25, (1.0, 1.0), (2.0, 1.0, 1.0)),
    (0.11, (0.7, 1.3), (3.1, 1.7, 2.2)),
    (0.5, (1.0, 1.0), (2.0, 1.0, 1.0)),
]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def seeded_state(rng, nx, ny, dtype, kind):
    """Seeded fields: Gaussian noise, noise plus flat runs (exercises the
    eps/eps ratio and the limiter's clamps), or a smooth bump."""
    if kind == 0:
        g = rng.standard_normal((ny, nx, 5))
    elif kind == 1:
        g = np.round(rng.standard_normal((ny, nx, 5)) * 2.0) / 2.0
        g[:, : nx // 2, :] = 0.25
    else:
        y, x = np.mgrid[0:ny, 0:nx]
        bump = np.exp(-((x - nx / 2.0) ** 2 + (y - ny / 3.0) ** 2) / (0.1 * nx * ny + 1))
        g = np.stack([bump * (k + 1) * 0.3 for k in range(5)], axis=-1)
    return np.ascontiguousarray(g.astype(dtype))


def tiny(dtype, out_name):
    ref = Reference(dtype)
    rng = np.random.default_rng(20240610 if dtype == np.float64 else 20240611)
    data = {}
    meta = []
    for si, (nx, ny) in enumerate(TINY_SHAPES):
        for ci, (dt, h, mat) in enumerate(TINY_CASES):
            kind = (si + ci) % 3
            g0 = seeded_state(rng, nx, ny, dtype, kind)
            key = f"s{si}c{ci}"
            meta.append({"key": key, "nx": nx, "ny": ny, "dt": dt, "h": list(h),
                         "mat": list(mat), "kind": kind})
            data[key + "_init"] = g0

            def fresh(stype=1, init=g0):
                t = ref.task(nx, ny, dt, h, mat, stype=stype)
                v = ref.view(t)
                v[...] = init if stype == 1 else np.moveaxis(init, -1, 0)
                return t

            t = fresh()
            ref.lib.le_step_x(ctypes.byref(t))
            data[key + "_x"] = ref.view(t).copy()
            ref.lib.le_free_task(ctypes.byref(t))

            t = fresh()
            ref.lib.le_step_y(ctypes.byref(t))
            data[key + "_y"] = ref.view(t).copy()
            ref.lib.le_free_task(ctypes.byref(t))

            t = fresh()
            ref.lib.le_step_y_cf(ctypes.byref(t), max(1, nx // 2))
            data[key + "_ycf"] = ref.view(t).copy()
            ref.lib.le_free_task(ctypes.byref(t))

            t = fresh()
            for s in range(10):
                ref.lib.le_step(ctypes.byref(t))
                if s in (0, 1, 9):
                    data[f"{key}_step{s + 1}"] = ref.view(t).copy()
            ref.lib.le_free_task(ctypes.byref(t))

            t = fresh(stype=0)
            for s in range(10):
                ref.lib.le_step_soa(ctypes.byref(t))
            data[key + "_soa10"] = np.moveaxis(ref.view(t), 0, -1).copy()
            ref.lib.le_free_task(ctypes.byref(t))
    np.savez_compressed(os.path.join(HERE, out_name), **data)
    with open(os.path.join(HERE, out_name.replace(".npz", ".json")), "w") as f:
        json.dump({"dtype": np.dtype(dtype).name, "cases": meta}, f, indent=1)


def config0():
    """512x512, dt = 0.5*h/c1 = 0.25, material (2, 1, 1), disk at (256, 256),
    r = 10, s = 1, 100 steps."""
    ref = Reference(np.float64)
    nx = ny = 512
    dt, h, mat = 0.25, (1.0, 1.0), (2.0, 1.0, 1.0)
    t = ref.task(nx, ny, dt, h, mat)
    ref.lib.le_set_ball(ctypes.byref(t), ref.abi.le_vec2(256.0, 256.0), 10.0, 1.0)
    g = ref.view(t)
    ball_nodes = int(np.count_nonzero(g[..., 2]))
    init_sha = sha(g)
    for _ in range(100):
        ref.lib.le_step(ctypes.byref(t))
    g = ref.view(t).copy()
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "c0.vtk")
        assert ref.lib.le_save_task(ctypes.byref(t), p.encode()) == 0
        vtk_sha = hashlib.sha256(open(p, "rb").read()).hexdigest()
    ref.lib.le_free_task(ctypes.byref(t))
    rng = np.random.default_rng(512)
    idx = rng.choice(nx * ny, size=1024, replace=False)
    flat = g.reshape(-1, 5)
    np.savez_compressed(os.path.join(HERE, "config0_samples_d.npz"), index=idx,
                        values=flat[idx])
    out = {
        "nx": nx, "ny": ny, "dt": dt, "h": list(h), "mat": list(mat),
        "ball": {"c": [256.0, 256.0], "r": 10.0, "s": 1.0, "nodes": ball_nodes,
                 "init_sha256": init_sha},
        "steps": 100,
        "state_sha256": sha(g),
        "field_sha256": {f: sha(g[..., k]) for k, f in
                         enumerate(("vx", "vy", "sxx", "sxy", "syy"))},
        "field_sum": [float(x).hex() for x in g.reshape(-1, 5).sum(axis=0)],
        "field_linf": [float(x).hex() for x in np.abs(g).reshape(-1, 5).max(axis=0)],
        "vtk_sha256": vtk_sha,
    }
    with open(os.path.join(HERE, "config0.json"), "w") as f:
        json.dump(out, f, indent=1)


def materials():
    rng = np.random.default_rng(77)
    rows = []
    for _ in range(16):
        c1 = float(rng.uniform(1.0, 4.0))
        c2 = c1 / float(rng.uniform(1.5, 2.2))
        rho = float(rng.uniform(1.0, 3.0))
        rows.append([c1, c2, rho])
    rows += [[2.0, 1.0, 1.5], [2.0, 1.0, 1.0]]
    out = {"inputs": [[x.hex() for x in r] for r in rows]}
    names = ("c1", "c2", "rho", "irhoc1", "irhoc2", "rhoc1", "rhoc2", "rhoc3")
    for dt in (np.float64, np.float32):
        ref = Reference(dt)
        res = []
        for r in rows:
            m = ref.material(*r)
            res.append([float(getattr(m, n)).hex() for n in names])
        out[np.dtype(dt).name] = res
    with open(os.path.join(HERE, "materials.json"), "w") as f:
        json.dump(out, f, indent=1)


def ball():
    cases = [((40, 30), (1.0, 1.0), (20.0, 15.0), 10.0, 1.0),
             ((40, 30), (0.7, 1.3), (13.3, 17.9), 6.5, -2.5),
             ((25, 64), (1.0, 0.5), (0.0, 0.0), 9.0, 0.75),
             ((33, 33), (1.0, 1.0), (16.0, 16.0), 4.0, 1.0)]
    out = []
    for dt in (np.float64, np.float32):
        ref = Reference(dt)
        for (nx, ny), h, c, r, s in cases:
            t = ref.task(nx, ny, 0.1, h, (2.0, 1.0, 1.0))
            ref.lib.le_set_ball(ctypes.byref(t), ref.abi.le_vec2(*c), r, s)
            g = ref.view(t)
            nz = np.argwhere(g[..., 2] != 0)  # (j, i) pairs
            out.append({"dtype": np.dtype(dt).name, "n": [nx, ny], "h": list(h), "c": list(c),
                        "r": r, "s": s, "nodes": nz.tolist(),
                        "sha256": sha(g)})
            ref.lib.le_free_task(ctypes.byref(t))
    with open(os.path.join(HERE, "ball.json"), "w") as f:
        json.dump(out, f)


def cli():
    out = {}
    for suffix in ("d", "f"):
        exe = os.path.join(ROOT, "oracle", "_ref", "simple_" + suffix)
        for args in (["64", "48", "20", "1"], ["100", "37", "7", "2"]):
            with tempfile.TemporaryDirectory() as td:
                r = subprocess.run([exe] + args, cwd=td, capture_output=True, text=True,
                                   check=True)
                fields = r.stdout.split()
                data = open(os.path.join(td, "result.vtk"), "rb").read()
                out[suffix + ":" + " ".join(args)] = {
                    "stdout_prefix": " ".join(fields[:3]),
                    "vtk_sha256": hashlib.sha256(data).hexdigest(),
                    "vtk_bytes": len(data)}
    with open(os.path.join(HERE, "cli.json"), "w") as f:
        json.dump(out, f, indent=1)


def tiny_vtk():
    ref = Reference(np.float64)
    rng = np.random.default_rng(5)
    t = ref.task(13, 7, 0.2, (0.5, 2.25), (2.0, 1.0, 1.0))
    ref.view(t)[...] = rng.standard_normal((7, 13, 5))
    np.save(os.path.join(HERE, "tiny_vtk_d_state.npy"), ref.view(t).copy())
    assert ref.lib.le_save_task(ctypes.byref(t), os.path.join(HERE, "tiny_vtk_d.vtk").encode()) == 0
    ref.lib.le_free_task(ctypes.byref(t))


if __name__ == "__main__":
    tiny(np.float64, "tiny_d.npz")
    tiny(np.float32, "tiny_f.npz")
    config0()
    materials()
    ball()
    cli()
    tiny_vtk()
    print("golden vectors written to", HERE)
