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


def mixed_instances(count=256, seed=3):
    """Mixed-size workload: sides log-uniform in [128, 2048] (independent),
    c1 ~ U[1,4], c2 = c1/U[1.5,2.2], rho ~ U[1,3], dt = 0.5*h/c1 with h = 1,
    disk radius U[4,32] inside the grid, sxx = syy = 1."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nx, ny = (int(round(np.exp(rng.uniform(np.log(128), np.log(2048))))) for _ in range(2))
        c1 = float(rng.uniform(1.0, 4.0))
        c2 = c1 / float(rng.uniform(1.5, 2.2))
        rho = float(rng.uniform(1.0, 3.0))
        r = float(rng.uniform(4.0, 32.0))
        cx, cy = float(rng.uniform(r, nx - r)), float(rng.uniform(r, ny - r))
        out.append(dict(nx=nx, ny=ny, c1=c1, c2=c2, rho=rho, dt=0.5 / c1, r=r, cx=cx, cy=cy))
    return out


def run_batch(le2d_amd, dtype, insts, steps, plain_steps):
    """Batched launch (autotuned, graph-replayed) vs one plain launch per
    instance per step, same instances.  Returns (result dict, batch)."""
    b = le2d_amd.Batch(dtype)
    for p in insts:
        b.add(p["nx"], p["ny"], p["dt"], (1.0, 1.0), (p["c1"], p["c2"], p["rho"]))
    b.commit()
    fb, seg, _ = b.autotune(2)  # on the zero state, which stays exactly zero
    for i, p in enumerate(insts):
        b.set_ball(i, (p["cx"], p["cy"]), p["r"], 1.0)
    nodes = sum(p["nx"] * p["ny"] for p in insts)
    ms_batch = b.time(steps)
    ms_plain = b.time(plain_steps, plain=True)
    b.wait()
    es = np.dtype(dtype).itemsize
    res = {"instances": len(insts), "nodes_per_step": nodes, "steps": steps,
           "shape": {"window_lanes": fb, "segment_rows": seg},
           "batched_Gnode_per_s": round(nodes / ms_batch / 1e6, 3),
           "plain_Gnode_per_s": round(nodes / ms_plain / 1e6, 3),
           "speedup_vs_plain": round(ms_plain / ms_batch, 3),
           "ms_per_step_batched": round(ms_batch, 4), "ms_per_step_plain": round(ms_plain, 4),
           "plain_steps_timed": plain_steps,
           "achieved_GBps_2sweep": round(nodes * 20 * es / ms_batch / 1e6, 1)}
    return res, b


def field_errors(a, ref):
    """Relative L2 and Linf of `a` against `ref` over all fields (FP64 sums)."""
    d = a.astype(np.float64) - ref
    return (float(np.sqrt(np.sum(d * d))), float(np.sqrt(np.sum(ref * ref))),
            float(np.max(np.abs(d))), float(np.max(np.abs(ref))))


def batch_report(le2d_amd, steps=500, plain_steps=50, count=256):
    insts = mixed_instances(count)
    out = {"workload": f"{count} instances, sides log-uniform [128, 2048], per-instance material "
                       f"and disk, {steps} steps"}
    r64, b64 = run_batch(le2d_amd, np.float64, insts, steps, plain_steps)
    r32, b32 = run_batch(le2d_amd, np.float32, insts, steps, plain_steps)
    # both batches advanced steps + plain_steps (+1 warm-up inside time()) from the disk
    e2n = e2d = einf = rinf = 0.0
    worst = 0.0
    for i in range(len(insts)):
        ref = b64.read(i)
        a = b32.read(i)
        n2, d2, ni, di = field_errors(a, ref)
        e2n += n2 * n2
        e2d += d2 * d2
        einf, rinf = max(einf, ni), max(rinf, di)
        worst = max(worst, n2 / d2 if d2 > 0 else 0.0)
    b64.close()
    b32.close()
    r32["l2_rel_vs_fp64"] = float(np.sqrt(e2n / e2d))
    r32["l2_rel_vs_fp64_worst_instance"] = worst
    r32["linf_rel_vs_fp64"] = einf / rinf
    out["fp32"] = r32
    out["fp64"] = r64
    return out


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
    ap.add_argument("--no-batch", action="store_true", help="skip the mixed-size batch leg")
    ap.add_argument("--batch-steps", type=int, default=500)
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

    batched = None
    if not args.no_batch:
        print("[bench] mixed-size batch", file=sys.stderr, flush=True)
        batched = batch_report(le2d_amd, steps=args.batch_steps)

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
        "batched": batched,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
