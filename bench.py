"""Throughput of the split-step elasticity sweeps on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 8192]
                    [--no-cpu] [--no-batch] [--no-tier] [--no-fp32]

Headline workload (`value`): the plain GPU sweep at scale.  An n x n grid
(default 8192^2), FP64, material c1 = 2, c2 = 1, rho = 1, h = 1,
dt = 0.5*h/c1 = 0.25, all fields zero except a disk of radius 64 cells
(sxx = syy = 1) at a seeded position in [64, n-64)^2.  A step is one x-sweep
plus one y-sweep over the whole grid, driven through the le_core.h C ABI
(le_step_n) with the fields resident in HBM.  W untimed steps, then exactly K
steps timed between two completion waits.

Further legs (each in its own key of the one JSON line printed):
  fp32     the same workload in FP32, its error against the FP64 run
  batched  256 instances of mixed sizes and materials: one persistent
           launch per step vs one launch per instance (FP32 and FP64)
  tier     16384^2, two materials split at the midline, 2000 steps, FP32 vs
           FP64 field error and energy drift
  serving  Poisson request stream through the serving loop (continuous
           batching under an HBM budget) vs the serial plain path: throughput
           and p50/p99 completion latency
  cpu_baseline  the reference's own simple/ path on the host cores
"""
import argparse
import ctypes
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E
METRIC = "Gnode-updates/s & % HBM peak, batched vs plain at stated error; p50/p99 latency"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def disk_centre(n, seed=1):
    rng = np.random.default_rng(seed)
    return tuple(float(v) for v in rng.integers(64, n - 64, size=2))


def plain_task(le2d_amd, n, dtype):
    t = le2d_amd.Task(n, n, 0.25, (1.0, 1.0), (2.0, 1.0, 1.0), dtype=dtype)
    t.set_ball(disk_centre(n), 64.0, 1.0)
    return t


def run_plain(le2d_amd, n, dtype, steps, warmup, profile_steps):
    """Times `steps` steps through le_step_n (two steps per pass); then times
    the two-step kernel, the single-step kernel and the two sweep kernels per
    launch with HIP events.  Every run of this function advances the task by
    warmup + steps + 4*profile_steps steps.  The two-step segment is tuned
    first (the state does not advance), so that no warm-up count leaves the
    tuning that le_step_n would do on its first call of 8 or more steps
    inside the timed region (LE2D_TUNE=0 keeps the built-in choice)."""
    t = plain_task(le2d_amd, n, dtype)
    if os.environ.get("LE2D_TUNE") != "0" and t.get_step_pairs() > 0:
        t.tune_step_pairs()
    t.step_n(max(1, warmup))
    t.wait()
    els = time_runs(t, steps, 1)
    ms_pair = t.time_pairs(profile_steps) if t.get_step_pairs() > 0 else None
    ms_step = t.time_steps(profile_steps)
    ms_x, ms_y = t.time_sweeps(profile_steps)
    return t, els, ms_pair, ms_step, ms_x, ms_y


def time_runs(t, steps, count):
    """Wall-clock seconds of `count` back-to-back runs of `steps` steps."""
    els = []
    for _ in range(count):
        t0 = time.perf_counter()
        t.step_n(steps)
        t.wait()
        els.append(time.perf_counter() - t0)
    return els


def sustained(t, n, steps, el, seconds):
    """Rate over at least `seconds` of back-to-back steps (an even count),
    run right after the timed steps: the clock the device holds under
    sustained load rather than in a short window.  The state keeps evolving
    (the wave spreads over more of the grid), so the shader clock and
    package power are sampled during the window."""
    k = max(steps, int(np.ceil(seconds / max(el / steps, 1e-6))))
    k += k & 1

    def run():
        t0 = time.perf_counter()
        t.step_n(k)
        t.wait()
        return time.perf_counter() - t0
    dt, sensors = sampled(run)
    out = {"steps": k, "seconds": round(dt, 3), "Gnode_per_s": round(n * n * k / dt / 1e9, 3)}
    out.update(sensors)
    return out


def gpu_sensors():
    """Shader clock (MHz) and package power (W) of the first GPU from sysfs,
    where the driver exposes them; {} otherwise."""
    import glob
    for d in sorted(glob.glob("/sys/class/drm/card*/device")):
        out = {}
        try:
            cur = [ln for ln in open(d + "/pp_dpm_sclk").read().splitlines()
                   if ln.strip().endswith("*")]
            if cur:
                out["sclk_MHz"] = float(cur[0].split()[1].lower().rstrip("mhz"))
        except (OSError, ValueError, IndexError):
            pass
        for h in glob.glob(d + "/hwmon/hwmon*"):
            for f in ("power1_input", "power1_average"):
                try:
                    out["power_W"] = int(open(os.path.join(h, f)).read()) / 1e6
                    break
                except (OSError, ValueError):
                    pass
        if out:
            return out
    return {}


def sampled(fn, period=0.02):
    """Run fn() while a thread samples gpu_sensors(); (result, medians)."""
    import threading
    samples, stop = [], [False]

    def loop():
        while not stop[0]:
            samples.append(gpu_sensors())
            time.sleep(period)
    th = threading.Thread(target=loop, daemon=True)
    th.start()
    try:
        res = fn()
    finally:
        stop[0] = True
        th.join()
    med = {}
    for k in ("sclk_MHz", "power_W"):
        v = [x[k] for x in samples if k in x]
        if v:
            med[k] = round(float(np.median(v)), 1)
    return res, med


def dense_report(le2d_amd, n, dtype, shape, passes=150, seed=5):
    """The same grid, the same launch shape, on the bench's disk field and
    with all five fields random (uniform in [-1, 1]): the kernels have no
    data-dependent path (the same instructions and cycles per launch), so
    any difference in rate comes from the clock the device holds; the shader
    clock and the power are sampled from sysfs during each timed run."""
    out = {"workload": f"{n}x{n}, {passes} two-step passes at the headline's launch shape "
                       f"(pair_seg, lanes, split = {shape}), on the disk field and on all five "
                       f"fields uniform in [-1, 1] (seed {seed})"}
    rng = np.random.default_rng(seed)
    for kind in ("disk", "dense"):
        with le2d_amd.Task(n, n, 0.25, (1.0, 1.0), (2.0, 1.0, 1.0), dtype=dtype) as t:
            if kind == "disk":
                t.set_ball(disk_centre(n), 64.0, 1.0)
            else:
                v = t.view()
                for j in range(0, n, 512):  # fill in slabs (bounded host temporaries)
                    v[j:j + 512] = rng.uniform(-1.0, 1.0, size=v[j:j + 512].shape)
                t.lib.le_mark_host_dirty(t.ref)
            t.set_step_pairs(shape[0])
            t.set_step_pair_window(shape[1])
            t.set_step_pair_split(shape[2])
            t.step_n(10)
            t.wait()
            ms, sens = sampled(lambda: t.time_pairs(passes))
        out[kind] = dict(Gnode_per_s=round(2 * n * n / ms / 1e6, 3), ms_per_pass=round(ms, 4),
                         **sens)
    out["dense_vs_disk"] = round(out["dense"]["Gnode_per_s"] / out["disk"]["Gnode_per_s"], 3)
    return out


def repeat_stats(n, steps, els):
    """Throughput of each back-to-back timed run: median, min and max."""
    v = sorted(n * n * steps / e / 1e9 for e in els)
    return {"runs": len(v), "median": round(float(np.median(v)), 3), "min": round(v[0], 3),
            "max": round(v[-1], 3)}


def measured_counters(n, dtype, kernel):
    """The committed rocprofv3 counter summary of `kernel` on this workload
    (profiles/traffic_<kernel>_<n>_<dtype>.json: HBM bytes per launch from
    FETCH_SIZE / WRITE_SIZE, VALU issue from the SQ counters), or {}."""
    p = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{n}_{dtype}.json")
    if not os.path.exists(p):
        return {}
    return json.load(open(p))


def roofline(n, es, ms_pair, ms_step, ms_x, ms_y, dtype_name, split=0):
    """Roofline of the launch that carries le_step_n: the two-step kernel
    (two node-updates per node per launch; k_step2_split when the tuner
    chose 64-lane windows on two wavefronts) when it is on, else the
    single-step kernel."""
    nodes = n * n
    pair_kernel = "k_step2_split" if split else "k_step2_xy"
    kernel, ms, steps = (pair_kernel, ms_pair, 2) if ms_pair else ("k_step_xy", ms_step, 1)
    # Algorithmic bytes per node-update: 2 sweeps x (5 reads + 5 writes) x
    # sizeof(real), times the node-updates per node one launch performs.
    alg = nodes * 20 * es * steps
    achieved = alg / (ms * 1e-3) / 1e9
    counters = measured_counters(n, dtype_name, kernel)
    traffic = counters.get("hbm_bytes_per_launch")
    # a fused launch moves each node's five values in and out once, however
    # many steps it takes: that is its compulsory traffic; `traffic` is what
    # rocprof counted
    compulsory = nodes * 10 * es
    actual = traffic / (ms * 1e-3) / 1e9 if traffic else None
    return {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": traffic,
            "traffic_GBps": round(actual, 1) if actual else None,
            "traffic_frac": round(actual / PEAK_HBM_GBS, 4) if actual else None,
            # the fraction of the VALU issue rate the kernel uses (SQ
            # counters of the committed profile): what bounds it, since its
            # actual traffic stays well below the HBM peak
            "valu_issue_frac": counters.get("valu_issue_frac"),
            "waves_per_simd": counters.get("waves_per_simd"),
            "valu_insts_per_node_update": counters.get("valu_insts_per_node_update"),
            "frac_basis": "achieved counts the two-sweep bytes of every node-update (160 B FP64, "
                          "80 B FP32); a two-step launch moves each node once for two steps, so "
                          "frac can exceed 1; traffic_frac (counted HBM bytes) and "
                          "valu_issue_frac are at most 1",
            "compulsory_bytes_per_launch": compulsory,
            "algorithmic_bytes_per_launch": alg, "steps_per_launch": steps,
            "ms_per_launch": {pair_kernel: round(ms_pair, 4) if ms_pair else None,
                              "k_step_xy": round(ms_step, 4), "k_sweep_x": round(ms_x, 4),
                              "k_sweep_y": round(ms_y, 4)},
            "two_kernel_GBps": {"k_sweep_x": round(nodes * 10 * es / ms_x / 1e6, 1),
                                "k_sweep_y": round(nodes * 10 * es / ms_y / 1e6, 1)}}


def rel_errors(norms):
    e2, r2, einf, rinf = norms
    return {"l2_rel": e2 / r2 if r2 > 0 else None, "linf_rel": einf / rinf if rinf > 0 else None}


def reference_case_report(le2d_amd, n=512, steps=100):
    """The reference driver's own case (512x512, h = 1, material (2, 1, 1),
    dt = 0.25, disk r = 10 at the centre, 100 steps): FP32 against FP64."""
    tasks = {}
    for dt in (np.float64, np.float32):
        t = le2d_amd.Task(n, n, 0.25, (1.0, 1.0), (2.0, 1.0, 1.0), dtype=dt)
        t.set_ball((float(n // 2), float(n // 2)), 10.0, 1.0)
        t.step_n(steps)
        tasks[dt] = t
    errs = rel_errors(le2d_amd.diff_norms(tasks[np.float32].device_view(),
                                          tasks[np.float64].device_view()))
    for t in tasks.values():
        t.close()
    return {"workload": f"{n}x{n}, disk r=10 at the centre, c1=2 c2=1 rho=1, dt=0.25, "
                        f"{steps} steps", "fp32_vs_fp64": errs}


def host_boundary_report(le2d_amd, n=8192, steps=100, trips=3):
    """The boundary's host-buffer path, PCIe included: the reference driver's
    pattern (state written into task->grid on the host, `steps` calls of
    le_step, the state read back) and a round trip per step (upload, one
    le_step, download)."""
    t = le2d_amd.Task(n, n, 0.25, (1.0, 1.0), (2.0, 1.0, 1.0))
    cx, cy = disk_centre(n)
    t.set_ball((cx, cy), 64.0, 1.0)
    t.step()
    t.sync_host()  # first upload and kernel load out of the timing
    t.lib.le_mark_host_dirty(t.ref)
    t0 = time.perf_counter()
    for _ in range(steps):
        t.step()
    t.sync_host()
    el = time.perf_counter() - t0
    t1 = time.perf_counter()
    for _ in range(trips):
        t.lib.le_mark_host_dirty(t.ref)
        t.step()
        t.sync_host()
    el2 = time.perf_counter() - t1
    # le_save_task after a step: the download and the legacy VTK of |v|
    saves = []
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "state.vtk")
        for _ in range(2):
            t.step()
            t.wait()
            t2 = time.perf_counter()
            if t.save(path) != 0:
                raise RuntimeError("le_save_task failed")
            saves.append(time.perf_counter() - t2)
            vtk_bytes = os.path.getsize(path)
    t.close()
    return {"workload": f"{n}x{n} FP64 through le_step on the host mirror",
            "driver_pattern": {"steps": steps, "seconds": round(el, 3),
                               "Gnode_per_s": round(n * n * steps / el / 1e9, 3)},
            "round_trip_per_step": {"ms_per_step": round(el2 / trips * 1e3, 2),
                                    "Gnode_per_s": round(n * n * trips / el2 / 1e9, 3)},
            "save_task": {"seconds": round(min(saves), 3), "vtk_bytes": vtk_bytes,
                          "what": "le_save_task after a step: download of the state and "
                                  "the VTK of |v| written to a local file (best of 2)"}}


def slab_report(le2d_amd, n=8192, parts=8, steps=400):
    """One GPU's share of the grid split into `parts` row slabs (the
    decomposition a multi-GPU run would use): the same path on an n x n/parts
    grid, FP64 and FP32."""
    out = {"workload": f"{n}x{n // parts} slab (1/{parts} of {n}x{n}), disk r=64, "
                       f"{steps} steps through le_step_n"}
    for dt, name in ((np.float64, "fp64"), (np.float32, "fp32")):
        with le2d_amd.Task(n, n // parts, 0.25, (1.0, 1.0), (2.0, 1.0, 1.0), dtype=dt) as t:
            t.set_ball((n / 2, n / parts / 2), 64.0, 1.0)
            t.step_n(10)
            t.wait()
            t0 = time.perf_counter()
            t.step_n(steps)
            t.wait()
            el = time.perf_counter() - t0
        out[name] = {"Gnode_per_s": round(n * (n // parts) * steps / el / 1e9, 3),
                     "ms_per_step": round(el / steps * 1e3, 4)}
    return out


def cpu_baseline(n, budget_s=12.0, single_budget_s=6.0):
    """Reference simple/ path (oracle/_ref, compiled from the reference's own
    sources) if present, else the oracle port; FP64, same grid and disk, a
    bounded number of steps on all host cores, and a shorter sample on one."""
    from oracle import Oracle, Reference
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cx, cy = disk_centre(n)
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
    # the same path on one core: the OpenMP runtime the CPU code is linked to
    single = None
    try:
        gomp = ctypes.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(1)
        t0 = time.perf_counter()
        k = 0
        while k < 20 and (k < 1 or time.perf_counter() - t0 < single_budget_s):
            step()
            k += 1
        el1 = time.perf_counter() - t0
        gomp.omp_set_num_threads(cores)
        single = {"value": n * n * k / el1 / 1e9, "cores": 1,
                  "sample": f"{n}x{n} FP64, {k} steps, {el1:.1f} s wall, one thread"}
    except OSError:
        pass
    done()
    return {"value": n * n * steps / el / 1e9, "unit": "Gnode-updates/s", "cores": cores,
            "kind": kind,
            "sample": f"{n}x{n} FP64, {steps} steps after 3 warm-up steps, "
                      f"{el:.1f} s wall, OMP_NUM_THREADS={cores}",
            "single_core": single}


# ---------------------------------------------------------------------------
# Mixed-size batch


def mixed_instances(count=256, seed=3):
    """Sides log-uniform in [128, 2048] (independent), c1 ~ U[1,4],
    c2 = c1/U[1.5,2.2], rho ~ U[1,3], dt = 0.5*h/c1 with h = 1, disk radius
    U[4,32] inside the grid, sxx = syy = 1."""
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


def run_batch(le2d_amd, dtype, insts, steps, plain_steps, shape=None):
    """Batched launch (autotuned, or the two-step form at a given (window,
    segment) shape; graph-replayed) vs one plain launch per instance per
    step, same instances."""
    name = np.dtype(dtype).name
    b = le2d_amd.Batch(dtype)
    for p in insts:
        b.add(p["nx"], p["ny"], p["dt"], (1.0, 1.0), (p["c1"], p["c2"], p["rho"]))
    b.commit()
    if shape is None:
        log(f"batch {name}: autotune")
        fb, seg, _ = b.autotune(2)  # on the zero state, which stays exactly zero
    else:
        fb, seg = shape
        b.set_shape(fb, seg, "pairs")
    for i, p in enumerate(insts):
        b.set_ball(i, (p["cx"], p["cy"]), p["r"], 1.0)
    nodes = sum(p["nx"] * p["ny"] for p in insts)
    b.step(4)  # warm-up: the step graph (two launches) is captured here, outside the timing
    b.wait()
    log(f"batch {name}: {steps} batched steps at {fb} lanes x {seg} rows")
    ms_batch = b.time(steps)
    log(f"batch {name}: {plain_steps} steps one launch per instance")
    ms_plain = b.time(plain_steps, plain=True)
    b.wait()
    es = np.dtype(dtype).itemsize
    res = {"instances": len(insts), "nodes_per_step": nodes, "steps": steps,
           "shape": {"window_lanes": fb, "segment_rows": seg,
                     "launch": {"tiles": "one workgroup per tile, one step",
                                "persistent": "persistent, one step",
                                "pairs": "one workgroup per tile, two steps",
                                "pairs_split": "one workgroup per tile, two steps, the two "
                                               "stages on two wavefronts"}[b.form()]},
           "batched_Gnode_per_s": round(nodes / ms_batch / 1e6, 3),
           "plain_Gnode_per_s": round(nodes / ms_plain / 1e6, 3),
           "speedup_vs_plain": round(ms_plain / ms_batch, 3),
           "ms_per_step_batched": round(ms_batch, 4), "ms_per_step_plain": round(ms_plain, 4),
           "plain_steps_timed": plain_steps,
           "achieved_GBps_2sweep": round(nodes * 20 * es / ms_batch / 1e6, 1)}
    return res, b


def batch_report(le2d_amd, steps=500, plain_steps=50, count=256, shape=None):
    insts = mixed_instances(count)
    out = {"workload": f"{count} instances, sides log-uniform [128, 2048], per-instance material "
                       f"and disk, {steps} steps timed batched, {plain_steps} per-instance"}
    r64, b64 = run_batch(le2d_amd, np.float64, insts, steps, plain_steps, shape)
    r32, b32 = run_batch(le2d_amd, np.float32, insts, steps, plain_steps, shape)
    # both advanced the same number of steps from the same disks: 4 warm-up,
    # the batched and the per-instance timed steps (autotuning ran on the
    # zero state before the disks were set)
    e2 = r2 = 0.0
    einf = rinf = 0.0
    per = []
    for i in range(len(insts)):
        n = le2d_amd.diff_norms(b32.device_view(i), b64.device_view(i))
        e2 += n[0] ** 2
        r2 += n[1] ** 2
        einf, rinf = max(einf, n[2]), max(rinf, n[3])
        per.append(n[0] / n[1] if n[1] > 0 else 0.0)
    b64.close()
    b32.close()
    per = np.array(per)
    w = int(np.argmax(per))
    r32["l2_rel_vs_fp64"] = float(np.sqrt(e2 / r2))
    r32["linf_rel_vs_fp64"] = einf / rinf
    r32["l2_rel_vs_fp64_per_instance"] = {
        "p50": float(np.percentile(per, 50)), "p90": float(np.percentile(per, 90)),
        "p99": float(np.percentile(per, 99)), "max": float(per.max()),
        "over_1e-5": int((per > 1e-5).sum())}
    r32["worst_instance"] = dict(index=w, l2_rel=float(per[w]),
                                 steps_from_disk=4 + steps + plain_steps, **insts[w])
    out["fp32"] = r32
    out["fp64"] = r64
    return out


# ---------------------------------------------------------------------------
# Precision tier with two materials


def tier_report(le2d_amd, n=16384, steps=2000, seed=7, every=250):
    """Two seeded materials split at x = n/2 and a disk of radius 64 placed
    256 cells left of the split, so that the P front (0.5 cells per step in
    the left material) reaches the interface after about 380 steps and the
    transmitted and reflected waves run for the rest of the 2000.  FP64 and
    FP32 runs from the same state; the energy is taken every `every` steps
    outside the timed segments.  After the run the FP64 state is read back
    once to show that the wave crossed: the largest |field| 200 cells right
    of the interface."""
    rng = np.random.default_rng(seed)
    mats = []
    for _ in range(2):
        c1 = float(rng.uniform(1.5, 3.0))
        mats.append((c1, c1 / float(rng.uniform(1.6, 2.0)), float(rng.uniform(1.0, 3.0))))
    dt = 0.5 / max(m[0] for m in mats)
    ids = np.zeros((n, n), np.uint8)
    ids[:, n // 2:] = 1
    cx, cy = float(n // 2 - 256), disk_centre(n, seed)[1]
    out = {"workload": f"{n}x{n}, two materials (c1, c2, rho) = "
                       f"{tuple(round(v, 4) for v in mats[0])} | "
                       f"{tuple(round(v, 4) for v in mats[1])} split at x = {n // 2}, "
                       f"dt = {dt:.4f}, disk r = 64 at ({cx:.0f}, {cy:.0f}), 256 cells left "
                       f"of the interface, {steps} steps"}
    tasks = {}
    for name, dtype in (("fp64", np.float64), ("fp32", np.float32)):
        log(f"tier {name}")
        t = le2d_amd.Task(n, n, dt, (1.0, 1.0), mats[0], dtype=dtype)
        t.set_ball((cx, cy), 64.0, 1.0)
        t.set_materials(mats, ids)
        e0 = t.energy()
        t.wait()
        curve, el, done = [1.0], 0.0, 0
        while done < steps:
            k = min(every, steps - done)
            t0 = time.perf_counter()
            t.step_n(k)
            t.wait()
            el += time.perf_counter() - t0
            done += k
            curve.append(round(t.energy() / e0, 6))
        e1 = t.energy()
        out[name] = {"Gnode_per_s": round(n * n * steps / el / 1e9, 3), "seconds": round(el, 3),
                     "energy_initial": e0, "energy_final": e1,
                     "energy_rel_drift": (e1 - e0) / e0,
                     f"energy_every_{every}_steps": curve,
                     "tiles": {"pair_seg": t.get_step_pairs(),
                               "pair_lanes": t.get_step_pair_window(),
                               "pair_split": t.get_step_pair_split()}}
        tasks[name] = t
    norms = le2d_amd.diff_norms(tasks["fp32"].device_view(), tasks["fp64"].device_view())
    out["fp32"].update({"vs_fp64_" + k: v for k, v in rel_errors(norms).items()})
    out["fp32"]["energy_vs_fp64_rel"] = (out["fp32"]["energy_final"] / out["fp64"]["energy_final"]
                                         - 1.0)
    g = tasks["fp64"].grid()
    out["fp64"]["peak_abs_200_cells_past_interface"] = float(np.abs(g[:, n // 2 + 200, :]).max())
    for t in tasks.values():
        t.close()
    return out


# ---------------------------------------------------------------------------
# Serving loop


def serve_requests(window_s, rate=200.0, seed=11):
    """Poisson arrivals at `rate` per second for `window_s` seconds; sides
    log-uniform in [256, 4096] (independent), steps U[200, 2000], 70% FP32 /
    30% FP64, per-request material (c1 ~ U[1,4], c2 = c1/U[1.5,2.2],
    rho ~ U[1,3], dt = 0.5/c1) and disk (r ~ U[4,32])."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t > window_s:
            return out
        nx, ny = (int(round(np.exp(rng.uniform(np.log(256), np.log(4096))))) for _ in range(2))
        c1 = float(rng.uniform(1.0, 4.0))
        r = float(rng.uniform(4.0, 32.0))
        out.append(dict(at=t, nx=nx, ny=ny, steps=int(rng.integers(200, 2001)),
                        dtype=np.float32 if rng.random() < 0.7 else np.float64, dt=0.5 / c1,
                        mat=(c1, c1 / float(rng.uniform(1.5, 2.2)), float(rng.uniform(1, 3))),
                        disk=(float(rng.uniform(r, nx - r)), float(rng.uniform(r, ny - r)), r,
                              1.0)))


def pct(v, q):
    return float(np.percentile(np.asarray(v), q)) if len(v) else None


def drain_logged(s, n, what, limit_s):
    """Drain the server, printing how many of its n requests are done every
    20 s; raises after limit_s seconds."""
    t0 = time.perf_counter()
    while not s.drained(20.0):
        if time.perf_counter() - t0 > limit_s:
            raise RuntimeError(f"{what}: not drained after {limit_s:.0f} s")
        log(f"{what}: {s.stats()['completed']} of {n} requests done")


def serial_sample(reqs, every):
    """A stratified sample of the requests for the serial baseline: within
    each precision, the requests ranked by work (nodes x steps) and every
    `every`-th of them taken, starting mid-stratum.  Returns, per sampled
    request, its index and the indices of the stratum it stands for."""
    out = []
    for dt in (np.float32, np.float64):
        ks = sorted((k for k, q in enumerate(reqs) if q["dtype"] == dt),
                    key=lambda k: reqs[k]["nx"] * reqs[k]["ny"] * reqs[k]["steps"])
        for a in range(0, len(ks), every):
            stratum = ks[a:a + every]
            out.append((stratum[len(stratum) // 2], stratum))
    return out


def serve_report(le2d_amd, window_s=60.0, budget_gb=200.0, slots=1024, chunk=16,
                 serial_every=20, plain_rates=None, split=0, serial=True, batch_rates=None):
    """Continuous batching of a Poisson request stream against the serial
    plain path.  The batched server receives the requests in real time.  The
    serial path runs a stratified sample (one request in `serial_every`,
    ranked by work within each precision) one at a time through the
    single-grid kernels; each sampled request's service rate stands for its
    stratum, and the serial queue is then replayed on the same arrivals.
    `plain_rates` (FP32, FP64 Gnode-updates/s of one large grid) gives the
    work-weighted ceiling of the mix; `batch_rates` (the mixed-size batch's
    FP32 and FP64 rates) the same ceiling at the rates of grids of the
    requests' sizes."""
    reqs = serve_requests(window_s)
    work = sum(q["nx"] * q["ny"] * q["steps"] for q in reqs)
    w32 = sum(q["nx"] * q["ny"] * q["steps"] for q in reqs if q["dtype"] == np.float32)
    out = {"workload": f"Poisson 200 req/s for {window_s:g} s ({len(reqs)} requests, "
                       f"{work / 1e9:.1f} Gnode-updates, {w32 / work:.2f} of it FP32), sides "
                       f"log-uniform [256, 4096], steps U[200, 2000], 70% FP32 / 30% FP64 "
                       f"requests, HBM budget {budget_gb:.0f} GB, {slots} slots, admission "
                       f"every {chunk} steps"}
    budget = int(budget_gb * (1 << 30))

    # the FP32 results of a few (the cheapest) FP32 requests are kept from the
    # mix and compared with the same requests run in FP64
    f32 = [k for k, q in enumerate(reqs) if q["dtype"] == np.float32]
    probe = sorted(f32, key=lambda k: reqs[k]["nx"] * reqs[k]["ny"] * reqs[k]["steps"])[:4]

    def submit(s, q, keep=False, dtype=None):
        return s.submit(q["nx"], q["ny"], q["steps"], dtype or q["dtype"], q["dt"], (1.0, 1.0),
                        q["mat"], q["disk"], keep=keep)

    # continuous batching, arrivals in real time
    log(f"serving {len(reqs)} requests over {window_s:g} s")
    with le2d_amd.Server(budget, slots) as s:
        s.set_shape(256, 16, chunk)
        s.set_pair_split(split)
        ids = []
        t0 = time.perf_counter()
        for k, q in enumerate(reqs):
            dt = q["at"] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids.append(submit(s, q, keep=k in probe))
        log(f"batched: all {len(reqs)} submitted")
        drain_logged(s, len(reqs), "batched", 300 + 5 * window_s)
        times = [s.times(i) for i in ids]
        st = s.stats()
        mixed = {k: s.result(ids[k]) for k in probe}
    errs = []
    with le2d_amd.Server(budget, slots) as s:
        twin = {k: submit(s, reqs[k], keep=True, dtype=np.float64) for k in probe}
        s.drain(300)
        for k in probe:
            ref = s.result(twin[k])
            d = mixed[k].astype(np.float64) - ref
            errs.append({"nx": reqs[k]["nx"], "ny": reqs[k]["ny"], "steps": reqs[k]["steps"],
                         "l2_rel": float(np.sqrt((d * d).sum() / (ref * ref).sum())),
                         "linf_rel": float(np.abs(d).max() / np.abs(ref).max())})
    out["fp32_vs_fp64"] = errs
    t_first = min(t[0] for t in times)
    span = max(t[2] for t in times) - t_first
    lat = [t[2] - t[0] for t in times]
    served = st["node_updates"] / span / 1e9
    out["batched"] = {"completed": st["completed"], "throughput_Gnode_per_s": round(served, 3),
                      "gpu_busy_Gnode_per_s": round(st["node_updates"] / st["busy_seconds"] / 1e9, 3),
                      "p50_latency_s": round(pct(lat, 50), 3), "p99_latency_s": round(pct(lat, 99), 3),
                      "makespan_s": round(span, 3), "peak_live": st["peak_live"],
                      "peak_GB": round(st["peak_bytes"] / 2 ** 30, 2)}
    if not serial:
        return out
    # serial plain path on a stratified sample: each sampled request's
    # service time (initialisation and steps, the gap between consecutive
    # completions of a one-slot server) scaled by work to its stratum
    sample = serial_sample(reqs, serial_every)
    log(f"serial baseline: {len(sample)} sampled requests")
    with le2d_amd.Server(budget, 1, plain=True) as s:
        sids = [submit(s, reqs[k]) for k, _ in sample]
        drain_logged(s, len(sample), "serial", 300 + 5 * window_s)
        done = [s.times(i)[2] for i in sids]
        svc = [done[0] - s.times(sids[0])[1]] + [b - a for a, b in zip(done, done[1:])]
    rate = {}
    for (k, stratum), d in zip(sample, svc):
        per = d / (reqs[k]["nx"] * reqs[k]["ny"] * reqs[k]["steps"])
        for m in stratum:
            rate[m] = per
    free, lat_s = 0.0, []
    for m, q in enumerate(reqs):
        start = max(q["at"], free)
        free = start + rate[m] * q["nx"] * q["ny"] * q["steps"]
        lat_s.append(free - q["at"])
    serial = work / (free - reqs[0]["at"]) / 1e9
    out["serial"] = {"throughput_Gnode_per_s": round(serial, 3),
                     "p50_latency_s": round(pct(lat_s, 50), 3),
                     "p99_latency_s": round(pct(lat_s, 99), 3), "makespan_s": round(free, 3),
                     "method": f"{len(sample)} requests run one at a time (one in {serial_every}, "
                               f"ranked by work within each precision); each one's service time "
                               f"per node-update stands for its stratum; the queue is replayed "
                               f"on the arrivals"}
    out["throughput_ratio_batched_vs_serial"] = round(served / serial, 3)
    if plain_rates:
        r32, r64 = plain_rates
        ceiling = work / 1e9 / (w32 / 1e9 / r32 + (work - w32) / 1e9 / r64)
        out["blended_ceiling_Gnode_per_s"] = round(ceiling, 3)
        out["ceiling_basis"] = (f"work-weighted: FP32 work at {r32:.1f} and FP64 work at "
                                f"{r64:.1f} Gnode-updates/s, the plain path on one 8192^2 grid")
        out["served_vs_ceiling"] = round(served / ceiling, 3)
    if batch_rates:
        r32, r64 = batch_rates
        ceiling = work / 1e9 / (w32 / 1e9 / r32 + (work - w32) / 1e9 / r64)
        out["blended_batch_ceiling_Gnode_per_s"] = round(ceiling, 3)
        out["batch_ceiling_basis"] = (f"work-weighted: FP32 work at {r32:.1f} and FP64 work at "
                                      f"{r64:.1f} Gnode-updates/s, the 256-instance mixed-size "
                                      f"batch (sides 128..2048)")
        out["served_vs_batch_ceiling"] = round(served / ceiling, 3)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--dtype", choices=["f64", "f32"], default="f64",
                    help="type of the headline run (the other type runs as a second leg)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the mixed-size batch leg")
    ap.add_argument("--no-tier", action="store_true", help="skip the two-material FP32 tier leg")
    ap.add_argument("--no-fp32", action="store_true", help="skip the second-type run")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop leg")
    ap.add_argument("--no-ref", action="store_true", help="skip the 512x512 reference-case leg")
    ap.add_argument("--no-host", action="store_true", help="skip the host-buffer (PCIe) leg")
    ap.add_argument("--no-slab", action="store_true", help="skip the one-GPU-slab leg")
    ap.add_argument("--serve-window", type=float, default=60.0,
                    help="seconds of Poisson arrivals in the serving leg")
    ap.add_argument("--serve-chunk", type=int, default=16,
                    help="longest run of steps between the serving loop's admission points")
    ap.add_argument("--serve-split", type=int, default=0, choices=(0, 1),
                    help="serving: 64-lane two-step windows on two wavefronts")
    ap.add_argument("--no-serial", action="store_true",
                    help="serving: skip the serial baseline")
    ap.add_argument("--serial-every", type=int, default=20,
                    help="serial baseline: one request in this many (stratified by work)")
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--batch-plain-steps", type=int, default=50,
                    help="steps of the one-launch-per-instance comparison loop")
    ap.add_argument("--batch-shape", default=None, metavar="LANES,ROWS",
                    help="two-step batch launches at this window and segment height "
                         "instead of autotuning (the same kernel in every profiling pass)")
    ap.add_argument("--tier-steps", type=int, default=2000)
    ap.add_argument("--sustain", type=float, default=2.0,
                    help="seconds of back-to-back steps after the timed ones (sustained rate)")
    ap.add_argument("--no-dense", action="store_true",
                    help="skip the all-fields-random leg")
    ap.add_argument("--repeats", type=int, default=5,
                    help="timed runs of --steps steps; value is the first, the median, "
                         "min and max of all are reported beside it")
    ap.add_argument("--profile-steps", type=int, default=20,
                    help="launches timed per kernel with HIP events for the roofline")
    args = ap.parse_args()
    args.repeats = max(1, args.repeats)
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU decomposition is a design, not a measured path", file=sys.stderr)
        sys.exit(2)

    import le2d_amd
    n = args.n
    main_dt = np.float64 if args.dtype == "f64" else np.float32
    es = np.dtype(main_dt).itemsize
    log(f"{n}x{n} {args.dtype}: {args.warmup} warm-up + {args.steps} timed steps")
    t, els, ms_pair, ms_step, ms_x, ms_y = run_plain(le2d_amd, n, main_dt, args.steps,
                                                     args.warmup, args.profile_steps)
    el = els[0]
    value = n * n * args.steps / el / 1e9
    roof = roofline(n, es, ms_pair, ms_step, ms_x, ms_y, args.dtype, t.get_step_pair_split())
    tiles = t.get_tiles() + (t.get_step_pairs(), t.get_step_pair_window(),
                             t.get_step_pair_split())
    other = None
    if not args.no_fp32:
        o_dt = np.float32 if main_dt == np.float64 else np.float64
        o_name = "f32" if o_dt == np.float32 else "f64"
        log(f"{n}x{n} {o_name}")
        t2, els2, mp2, ms2, mx2, my2 = run_plain(le2d_amd, n, o_dt, args.steps, args.warmup,
                                                 args.profile_steps)
        el2 = els2[0]
        lo, hi = (t2, t) if o_dt == np.float32 else (t, t2)
        errs = rel_errors(le2d_amd.diff_norms(lo.device_view(), hi.device_view()))
        other = {"dtype": o_name, "value": round(n * n * args.steps / el2 / 1e9, 3),
                 "ms_per_step": round(el2 / args.steps * 1e3, 4),
                 "roofline": roofline(n, np.dtype(o_dt).itemsize, mp2, ms2, mx2, my2, o_name,
                                      t2.get_step_pair_split()),
                 "tiles": {"pair_seg": t2.get_step_pairs(),
                           "pair_lanes": t2.get_step_pair_window(),
                           "pair_split": t2.get_step_pair_split()},
                 "fp32_vs_fp64_after_steps": args.warmup + args.steps + 4 * args.profile_steps,
                 "fp32_vs_fp64": errs}
    # The headline type's further timed runs and its sustained window follow
    # its first run directly (the other type's first run, which the error
    # comparison above needs at the same step count, is the only work in
    # between); then the other type's, then the dense-field legs.
    els += time_runs(t, args.steps, args.repeats - 1)
    log(f"{n}x{n} {args.dtype}: sustained {args.sustain:g} s")
    sus = sustained(t, n, args.steps, el, args.sustain)
    shape = (t.get_step_pairs(), t.get_step_pair_window(), t.get_step_pair_split())
    t.close()
    if other is not None:
        other["repeats"] = repeat_stats(n, args.steps,
                                        els2 + time_runs(t2, args.steps, args.repeats - 1))
        other["sustained"] = sustained(t2, n, args.steps, el2, args.sustain)
        shape2 = (t2.get_step_pairs(), t2.get_step_pair_window(), t2.get_step_pair_split())
        t2.close()
    dense = None
    if not args.no_dense:
        log(f"{n}x{n} {args.dtype} dense field")
        dense = dense_report(le2d_amd, n, main_dt, shape)
        if other is not None:
            log(f"{n}x{n} {other['dtype']} dense field")
            other["dense_field"] = dense_report(le2d_amd, n,
                                                np.float32 if other["dtype"] == "f32"
                                                else np.float64, shape2)
    copy_bytes = n * n * 5 * es
    copy_gbs = 2 * copy_bytes / le2d_amd.copy_probe(copy_bytes, 10) / 1e6

    batched = None
    if not args.no_batch:
        log("mixed-size batch")
        shape = tuple(int(v) for v in args.batch_shape.split(",")) if args.batch_shape else None
        batched = batch_report(le2d_amd, steps=args.batch_steps,
                               plain_steps=args.batch_plain_steps, shape=shape)
    ref_case = None if args.no_ref else reference_case_report(le2d_amd)
    host = None if args.no_host else host_boundary_report(le2d_amd, n)
    slab = None if args.no_slab else slab_report(le2d_amd, n)
    tier = None
    if not args.no_tier:
        tier = tier_report(le2d_amd, steps=args.tier_steps)
    serving = None
    if not args.no_serve:
        rates = None
        if other is not None:
            r64, r32 = (value, other["value"]) if main_dt == np.float64 else (other["value"], value)
            rates = (r32, r64)
        serving = serve_report(le2d_amd, window_s=args.serve_window, chunk=args.serve_chunk,
                               serial_every=args.serial_every, plain_rates=rates,
                               split=args.serve_split, serial=not args.no_serial,
                               batch_rates=(batched["fp32"]["batched_Gnode_per_s"],
                                            batched["fp64"]["batched_Gnode_per_s"])
                               if batched else None)
    cpu = None
    if not args.no_cpu:
        log("CPU baseline")
        cpu = cpu_baseline(n)

    cx, cy = disk_centre(n)
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gnode-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": args.dtype, "data": "synthetic",
        "config": {"workload": f"plain split-step sweep, {n}x{n} grid, disk r=64 at "
                               f"({cx:.0f},{cy:.0f}), c1=2 c2=1 rho=1, dt=0.25",
                   "nx": n, "ny": n,
                   "tiles": dict(zip(("x_rows", "y_seg", "y_block", "xy_seg", "xy_block",
                                      "pair_seg", "pair_lanes", "pair_split"), tiles))},
        "repeats": repeat_stats(n, args.steps, els),
        "sustained": sus,
        "dense_field": dense,
        "hbm_copy_GBps": round(copy_gbs, 1),
        "roofline": roof,
        "cpu_baseline": cpu,
        ("fp32" if main_dt == np.float64 else "fp64"): other,
        "reference_case": ref_case,
        "host_boundary": host,
        "slab": slab,
        "batched": batched,
        "tier": tier,
        "serving": serving,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
