"""Throughput of the split-step elasticity sweeps on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--dtype f64|f32] [--n 8192]

Workload (the plain GPU sweep at scale): an n x n grid (default 8192^2),
material c1 = 2, c2 = 1, rho = 1, h = 1, dt = 0.5*h/c1 = 0.25, all fields
zero except a disk of radius 64 cells (sxx = syy = 1) at a seeded position in
[64, n-64)^2.  A step is one x-sweep plus one y-sweep over the whole grid,
driven through the le_core.h C ABI (le_step_n) with the fields resident in
HBM.  W untimed steps, then exactly K steps timed between two completion
waits.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E
METRIC = "Gnode-updates/s & % HBM peak, batched vs plain at stated error; p50/p99 latency"


def make_task(le2d_amd, n, dtype, seed=1):
    rng = np.random.default_rng(seed)
    cx, cy = (float(v) for v in rng.integers(64, n - 64, size=2))
    t = le2d_amd.Task(n, n, 0.25, (1.0, 1.0), (2.0, 1.0, 1.0), dtype=dtype)
    t.set_ball((cx, cy), 64.0, 1.0)
    return t, (cx, cy)


def cpu_baseline(n, budget_s=12.0):
    """Reference simple/ path (oracle/_ref, compiled from the reference's own
    sources) if present, else the oracle port; FP64, same grid and disk, a
    bounded number of steps on all host cores."""
    from oracle import Oracle, Reference
    import ctypes
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rng = np.random.default_rng(1)
    cx, cy = (float(v) for v in rng.integers(64, n - 64, size=2))
    if Reference.available(np.float64):
        ref = Reference(np.float64)
        t = ref.task(n, n, 0.25, (1.0, 1.0), (2.0, 1.0, 1.0))
        ref.lib.le_set_ball(ctypes.byref(t), ref.abi.le_vec2(cx, cy), 64.0, 1.0)
        step = lambda: ref.lib.le_step(ctypes.byref(t))  # noqa: E731
        kind, done = "reference", lambda: ref.lib.le_free_task(ctypes.byref(t))  # noqa: E731
    else:
        o = Oracle(np.float64)
        g = np.zeros((n, n, 5))
        o.set_ball(g, (1.0, 1.0), (cx, cy), 64.0, 1.0)
        step = lambda: o.run(g, 0.25, (1.0, 1.0), (2.0, 1.0, 1.0), 1)  # noqa: E731
        kind, done = "port", lambda: None  # noqa: E731
    step()  # first touch of the host pages
    t0 = time.perf_counter()
    step()
    step()
    one = (time.perf_counter() - t0) / 2
    steps = max(1, min(200, int(budget_s / max(one, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    el = time.perf_counter() - t0
    done()
    return {"value": n * n * steps / el / 1e9, "unit": "Gnode-updates/s", "cores": cores,
            "kind": kind,
            "sample": f"{n}x{n} FP64, {steps} steps after 3 warm-up steps, "
                      f"{el:.1f} s wall, OMP_NUM_THREADS={cores}"}


def measured_traffic(n, dtype):
    """HBM bytes per k_step_xy launch from the committed rocprofv3 counter
    summary for this workload (profiles/traffic_k_step_xy_<n>_<dtype>.json),
    or None when no such measurement exists."""
    p = os.path.join(ROOT, "profiles", f"traffic_k_step_xy_{n}_{dtype}.json")
    if not os.path.exists(p):
        return None
    return json.load(open(p)).get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--dtype", choices=["f64", "f32"], default="f64")
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--profile-steps", type=int, default=20,
                    help="steps timed per sweep with HIP events for the roofline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU decomposition is a design, not a measured path", file=sys.stderr)
        sys.exit(2)

    import le2d_amd
    dtype = np.float64 if args.dtype == "f64" else np.float32
    es = np.dtype(dtype).itemsize
    n = args.n
    print(f"[bench] {n}x{n} {args.dtype}: building task", file=sys.stderr, flush=True)
    t, centre = make_task(le2d_amd, n, dtype)
    t.step_n(max(1, args.warmup))  # uploads, then warms up
    t.wait()
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    t.wait()
    t0 = time.perf_counter()
    t.step_n(args.steps)
    t.wait()
    el = time.perf_counter() - t0
    nodes = n * n
    value = nodes * args.steps / el / 1e9

    # Per-launch durations with HIP events on the task's stream.  le_step runs
    # one kernel per step (k_step_xy); the separate sweep kernels (le_step_x,
    # le_step_y) are timed too for reference.
    ms_step = t.time_steps(args.profile_steps)
    ms_x, ms_y = t.time_sweeps(args.profile_steps)
    # Algorithmic bytes per node-update: 2 sweeps x (5 reads + 5 writes) x
    # sizeof(real); one k_step_xy launch advances every node by one step.
    alg_bytes = nodes * 20 * es
    achieved = alg_bytes / (ms_step * 1e-3) / 1e9
    traffic = measured_traffic(n, args.dtype)
    roofline = {"bound": "hbm", "kernel": "k_step_xy", "achieved": round(achieved, 1),
                "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                "traffic": traffic,
                "algorithmic_bytes_per_launch": alg_bytes,
                "ms_per_launch": {"k_step_xy": round(ms_step, 4), "k_sweep_x": round(ms_x, 4),
                                  "k_sweep_y": round(ms_y, 4)},
                "two_kernel_GBps": {"k_sweep_x": round(nodes * 10 * es / ms_x / 1e6, 1),
                                    "k_sweep_y": round(nodes * 10 * es / ms_y / 1e6, 1)}}
    tiles = t.get_tiles()
    t.close()
    copy_ms = le2d_amd.copy_probe(nodes * 5 * es, 10)
    copy_gbs = 2 * nodes * 5 * es / copy_ms / 1e6

    cpu = None
    if not args.no_cpu:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(n)

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gnode-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if dtype == np.float64 else "f32", "data": "synthetic",
        "config": {"workload": f"plain split-step sweep, {n}x{n} grid, disk r=64 at "
                               f"({centre[0]:.0f},{centre[1]:.0f}), c1=2 c2=1 rho=1, dt=0.25",
                   "nx": n, "ny": n, "tiles": {"x_rows": tiles[0], "y_seg": tiles[1],
                                               "y_block": tiles[2], "xy_seg": tiles[3],
                                               "xy_block": tiles[4]}},
        "hbm_copy_GBps": round(copy_gbs, 1),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
