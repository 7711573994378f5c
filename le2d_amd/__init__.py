"""le2d_amd -- MI355X split-step elasticity solver behind the le_core.h C ABI.

The product is the pair of shared libraries built from le2d_amd/csrc
(lib/liblecore_d.so with real = double, lib/liblecore_f.so with real =
float).  This module is a thin ctypes front end over that ABI for Python
callers, tests and the benchmark; every numeric operation runs in the HIP
library.  There is no CPU fallback: loading fails if the library was not
built, and stepping aborts inside the library if no GPU is present.
"""
import ctypes
import os

import numpy as np

from .abi import ABI_D, ABI_F, ST_AOS, ST_SOA, NODE_FIELDS  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
# LE2D_LIBDIR selects another build of the same libraries (the host-sanitizer
# build of le2d_amd/csrc `make san` lives in le2d_amd/lib_san).
LIBDIR = os.environ.get("LE2D_LIBDIR") or os.path.join(HERE, "lib")

SYNC_WAIT, SYNC_ASYNC, SYNC_COHERENT = 0, 1, 2
# LE2D_ERR_* status codes of include/le2d.h
ERR_ARG, ERR_DEVICE, ERR_NOMEM, ERR_STATE, ERR_BUDGET, ERR_BUSY = 1, 2, 3, 4, 5, 6

# Symbols declared in include/le2d.h (extensions), beyond simple/le_core.h.
_EXT = {
    "le2d_last_error": ([], ctypes.c_char_p),
    "le2d_set_sync_mode": (["T", ctypes.c_int], ctypes.c_int),
    "le2d_wait": (["T"], ctypes.c_int),
    "le_sync_host": (["T"], ctypes.c_int),
    "le_mark_host_dirty": (["T"], ctypes.c_int),
    "le_step_n": (["T", ctypes.c_int], ctypes.c_int),
    "le2d_set_tiles": (["T"] + [ctypes.c_int] * 5, ctypes.c_int),
    "le2d_get_tiles": (["T"] + [ctypes.POINTER(ctypes.c_int)] * 5, ctypes.c_int),
    "le2d_time_sweeps": (["T", ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                          ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "le2d_time_steps": (["T", ctypes.c_int, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "le2d_set_step_pairs": (["T", ctypes.c_int], ctypes.c_int),
    "le2d_get_step_pairs": (["T", ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "le2d_set_step_pair_window": (["T", ctypes.c_int], ctypes.c_int),
    "le2d_get_step_pair_window": (["T", ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "le2d_set_step_pair_split": (["T", ctypes.c_int], ctypes.c_int),
    "le2d_get_step_pair_split": (["T", ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "le2d_tune_step_pairs": (["T", ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "le2d_time_pairs": (["T", ctypes.c_int, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "le2d_copy_probe": ([ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
                        ctypes.c_int),
    "le2d_device_bytes": (["T"], ctypes.c_size_t),
    "le2d_dump": (["T", ctypes.c_char_p], ctypes.c_int),
    "le2d_set_material_plane": (["T", ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p],
                                ctypes.c_int),
    "le2d_energy": (["T", ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "le2d_task_view": (["T", ctypes.c_void_p], ctypes.c_int),
    "le2d_restore": (["T", ctypes.c_char_p], ctypes.c_int),
}

class le2d_view(ctypes.Structure):
    _fields_ = [("base", ctypes.c_void_p), ("elem", ctypes.c_int), ("nx", ctypes.c_int),
                ("ny", ctypes.c_int), ("pitch", ctypes.c_long), ("plane", ctypes.c_long)]


class le2d_request(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("real_size", ctypes.c_int),
                ("steps", ctypes.c_int)] + [(n, ctypes.c_double) for n in
                                            ("dt", "hx", "hy", "c1", "c2", "rho", "cx", "cy",
                                             "r", "s")] + [("keep_result", ctypes.c_int)]


class le2d_server_stats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_longlong), ("queued", ctypes.c_longlong),
                ("live", ctypes.c_int), ("peak_live", ctypes.c_int),
                ("node_updates", ctypes.c_double), ("busy_seconds", ctypes.c_double),
                ("bytes_in_use", ctypes.c_size_t), ("peak_bytes", ctypes.c_size_t),
                ("cancelled", ctypes.c_longlong), ("rejected", ctypes.c_longlong)]


# include/le2d.h device allocator hook: int alloc(void **ptr, size_t bytes,
# void *ctx) and void release(void *ptr, void *ctx)
ALLOC_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t,
                            ctypes.c_void_p)
FREE_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p)


# include/le2d.h batched-instance API (no le_task argument).
_VP = ctypes.c_void_p
_IP = ctypes.POINTER(ctypes.c_int)
_DP = ctypes.POINTER(ctypes.c_double)
_BATCH = {
    "le2d_batch_create": ([ctypes.c_int], _VP),
    "le2d_batch_destroy": ([_VP], None),
    "le2d_batch_add": ([_VP, ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 6, ctypes.c_int),
    "le2d_batch_count": ([_VP], ctypes.c_int),
    "le2d_batch_commit": ([_VP], ctypes.c_int),
    "le2d_batch_device_bytes": ([_VP], ctypes.c_size_t),
    "le2d_batch_set_ball": ([_VP, ctypes.c_int] + [ctypes.c_double] * 4, ctypes.c_int),
    "le2d_batch_load": ([_VP, ctypes.c_int, _VP], ctypes.c_int),
    "le2d_batch_read": ([_VP, ctypes.c_int, _VP], ctypes.c_int),
    "le2d_batch_step": ([_VP, ctypes.c_int], ctypes.c_int),
    "le2d_batch_step_plain": ([_VP, ctypes.c_int], ctypes.c_int),
    "le2d_batch_wait": ([_VP], ctypes.c_int),
    "le2d_batch_set_shape": ([_VP, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "le2d_batch_set_launch": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "le2d_batch_get_launch": ([_VP, _IP], ctypes.c_int),
    "le2d_batch_get_shape": ([_VP, _IP, _IP, _IP, _IP], ctypes.c_int),
    "le2d_batch_time": ([_VP, ctypes.c_int, ctypes.c_int, _DP], ctypes.c_int),
    "le2d_batch_autotune": ([_VP, ctypes.c_int, _IP, _IP, _DP], ctypes.c_int),
    "le2d_batch_view": ([_VP, ctypes.c_int, ctypes.POINTER(le2d_view)], ctypes.c_int),
    "le2d_diff_norms": ([ctypes.POINTER(le2d_view), ctypes.POINTER(le2d_view), _DP],
                        ctypes.c_int),
    "le2d_diff_norms_noalloc": ([ctypes.POINTER(le2d_view), ctypes.POINTER(le2d_view), _DP],
                                ctypes.c_int),
    "le2d_server_create": ([ctypes.c_size_t, ctypes.c_int, ctypes.c_int], _VP),
    "le2d_server_set_shape": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "le2d_server_set_pairs": ([_VP, ctypes.c_int], ctypes.c_int),
    "le2d_server_set_pair_window": ([_VP, ctypes.c_int], ctypes.c_int),
    "le2d_server_set_pair_split": ([_VP, ctypes.c_int], ctypes.c_int),
    "le2d_server_set_admission": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int],
                                  ctypes.c_int),
    "le2d_server_submit": ([_VP, ctypes.POINTER(le2d_request)], ctypes.c_longlong),
    "le2d_server_cancel": ([_VP, ctypes.c_longlong], ctypes.c_int),
    "le2d_set_device_allocator": ([_VP, _VP, _VP], ctypes.c_int),
    "le2d_server_wait": ([_VP, ctypes.c_longlong, ctypes.c_double], ctypes.c_int),
    "le2d_server_drain": ([_VP, ctypes.c_double], ctypes.c_int),
    "le2d_server_result": ([_VP, ctypes.c_longlong, _VP], ctypes.c_int),
    "le2d_server_times": ([_VP, ctypes.c_longlong, _DP, _DP, _DP], ctypes.c_int),
    "le2d_server_get_stats": ([_VP, ctypes.POINTER(le2d_server_stats)], ctypes.c_int),
    "le2d_server_destroy": ([_VP], None),
    "le2d_limiter": ([], ctypes.c_int),
}

_libs = {}

# include/le2d.h LE2D_LIMITER_*: the flux limiter a library was built with.
# superbee (the reference's choice) is in LIBDIR; `make limiters` builds the
# minmod and CIR libraries into lib_minmod/ and lib_cir/.
LIMITERS = {"superbee": 0, "minmod": 1, "cir": 2}


def lib_path(dtype=np.float64, limiter="superbee"):
    if limiter not in LIMITERS:
        raise ValueError(f"limiter {limiter!r}: superbee, minmod or cir")
    name = "liblecore_d.so" if np.dtype(dtype) == np.float64 else "liblecore_f.so"
    base = LIBDIR if limiter == "superbee" else os.path.join(HERE, "lib_" + limiter)
    return os.path.join(base, name)


def load(dtype=np.float64, limiter="superbee"):
    """Load (once) the C-ABI library for `dtype` built with `limiter` and bind
    its entry points."""
    dtype = np.dtype(dtype)
    if (dtype, limiter) in _libs:
        return _libs[(dtype, limiter)]
    path = lib_path(dtype, limiter)
    if not os.path.exists(path):
        raise RuntimeError(f"le2d_amd: {path} is not built; run __graft_entry__.build() "
                           "(make -C le2d_amd/csrc all limiters)")
    lib = ctypes.CDLL(path)
    abi = ABI_D if dtype == np.float64 else ABI_F
    abi.bind(lib)
    PT = ctypes.POINTER(abi.le_task)
    for name, (args, res) in _EXT.items():
        fn = getattr(lib, name)
        fn.argtypes = [PT if a == "T" else a for a in args]
        fn.restype = res
    for name, (args, res) in _BATCH.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if lib.le2d_limiter() != LIMITERS[limiter]:
        raise RuntimeError(f"le2d_amd: {path} was not built with the {limiter} limiter")
    lib.abi = abi
    _libs[(dtype, limiter)] = lib
    return lib


class LE2DError(RuntimeError):
    pass


class BusyError(LE2DError):
    """The server's queue is at its limit (LE2D_ERR_BUSY): retry later."""


def set_device_allocator(alloc=None, release=None, dtype=np.float64):
    """Install device allocation callbacks (ALLOC_FN, FREE_FN instances) in
    the library of `dtype`, or restore hipMalloc / hipFree with None.  The
    caller keeps the callback objects alive while any allocation they made
    is in use."""
    lib = load(dtype)
    a = ctypes.cast(alloc, ctypes.c_void_p) if alloc is not None else None
    r = ctypes.cast(release, ctypes.c_void_p) if release is not None else None
    rc = lib.le2d_set_device_allocator(a, r, None)
    if rc != 0:
        raise LE2DError(f"set_device_allocator: {rc} {lib.le2d_last_error().decode()}")


def copy_probe(nbytes, iters=20, dtype=np.float64):
    """Mean ms of one device-to-device copy of `nbytes` (achievable HBM rate)."""
    lib = load(dtype)
    ms = ctypes.c_double()
    rc = lib.le2d_copy_probe(nbytes, iters, ctypes.byref(ms))
    if rc != 0:
        raise LE2DError(f"copy_probe: {rc} {lib.le2d_last_error().decode()}")
    return ms.value


def diff_norms(a, ref):
    """(||a-ref||_2, ||ref||_2, max|a-ref|, max|ref|) of two device views,
    FP64 on the device."""
    lib = load(np.float64)
    out = (ctypes.c_double * 4)()
    rc = lib.le2d_diff_norms(ctypes.byref(a), ctypes.byref(ref), out)
    if rc != 0:
        raise LE2DError(f"diff_norms: {rc} {lib.le2d_last_error().decode()}")
    return tuple(out)


class Task:
    """One le_task owned through the C ABI.  `grid()` returns a numpy view of
    the host mirror, (ny, nx, 5) for AOS or (5, ny, nx) for SOA, after making
    it current."""

    def __init__(self, nx, ny, dt, h=(1.0, 1.0), mat=(2.0, 1.0, 1.0), stype=ST_AOS,
                 dtype=np.float64, limiter="superbee"):
        self.dtype = np.dtype(dtype)
        self.lib = load(self.dtype, limiter)
        self.abi = self.lib.abi
        self.mat = self.abi.le_material()
        self.lib.le_init_material(mat[0], mat[1], mat[2], ctypes.byref(self.mat))
        self.t = self.abi.le_task()
        self.lib.le_init_task(ctypes.byref(self.t), dt, self.abi.le_vec2(*h), self.mat,
                              self.abi.le_point2(nx, ny), stype)
        self.nx, self.ny, self.stype = nx, ny, stype
        self._open = True

    @property
    def ref(self):
        return ctypes.byref(self.t)

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.le2d_last_error()
            raise LE2DError(f"{what}: {rc} {msg.decode() if msg else ''}")

    def view(self):
        """Numpy view of the host mirror without synchronising."""
        n = self.nx * self.ny * 5
        flat = np.ctypeslib.as_array(self.t.grid, shape=(n,))
        if self.stype == ST_AOS:
            return flat.reshape(self.ny, self.nx, 5)
        return flat.reshape(5, self.ny, self.nx)

    def grid(self):
        self.sync_host()
        return self.view()

    def load(self, state):
        """Write `state` (same shape as view()) into the mirror and mark it dirty."""
        self.sync_host()
        self.view()[...] = state
        self._check(self.lib.le_mark_host_dirty(self.ref), "mark_host_dirty")

    def set_ball(self, c, r, s):
        self.lib.le_set_ball(self.ref, self.abi.le_vec2(*c), r, s)

    def step(self):
        (self.lib.le_step if self.stype == ST_AOS else self.lib.le_step_soa)(self.ref)

    def step_n(self, n):
        self._check(self.lib.le_step_n(self.ref, n), "le_step_n")

    def step_x(self):
        (self.lib.le_step_x if self.stype == ST_AOS else self.lib.le_soa_step_x)(self.ref)

    def step_y(self):
        (self.lib.le_step_y if self.stype == ST_AOS else self.lib.le_soa_step_y)(self.ref)

    def step_y_cf(self, cfs):
        self.lib.le_step_y_cf(self.ref, cfs)

    def step_cf(self, cfs):
        self.lib.le_step_cf(self.ref, cfs)

    def sync_host(self):
        self._check(self.lib.le_sync_host(self.ref), "le_sync_host")

    def wait(self):
        self._check(self.lib.le2d_wait(self.ref), "le2d_wait")

    def set_sync_mode(self, mode):
        self._check(self.lib.le2d_set_sync_mode(self.ref, mode), "set_sync_mode")

    def set_tiles(self, x_rows=0, y_seg=0, y_block=0, xy_seg=0, xy_block=0):
        self._check(self.lib.le2d_set_tiles(self.ref, x_rows, y_seg, y_block, xy_seg, xy_block),
                    "set_tiles")

    def get_tiles(self):
        v = [ctypes.c_int() for _ in range(5)]
        self._check(self.lib.le2d_get_tiles(self.ref, *[ctypes.byref(x) for x in v]),
                    "get_tiles")
        return tuple(x.value for x in v)

    def time_sweeps(self, steps):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.le2d_time_sweeps(self.ref, steps, ctypes.byref(a),
                                              ctypes.byref(b)), "time_sweeps")
        return a.value, b.value

    def time_steps(self, steps):
        a = ctypes.c_double()
        self._check(self.lib.le2d_time_steps(self.ref, steps, ctypes.byref(a)), "time_steps")
        return a.value

    def set_step_pairs(self, seg):
        """Rows per segment of the two-step pass used by step_n (0: off)."""
        self._check(self.lib.le2d_set_step_pairs(self.ref, seg), "set_step_pairs")

    def set_step_pair_window(self, lanes):
        """Lanes per two-step workgroup: 64, 128 or 256 (writes lanes - 8 columns)."""
        self._check(self.lib.le2d_set_step_pair_window(self.ref, lanes), "set_step_pair_window")

    def get_step_pair_window(self):
        v = ctypes.c_int()
        self._check(self.lib.le2d_get_step_pair_window(self.ref, ctypes.byref(v)),
                    "get_step_pair_window")
        return v.value

    def set_step_pair_split(self, on):
        """64-lane windows: the two stages on two wavefronts (1) or one (0)."""
        self._check(self.lib.le2d_set_step_pair_split(self.ref, int(on)), "set_step_pair_split")

    def get_step_pair_split(self):
        v = ctypes.c_int()
        self._check(self.lib.le2d_get_step_pair_split(self.ref, ctypes.byref(v)),
                    "get_step_pair_split")
        return v.value

    def tune_step_pairs(self):
        """Time the two-step pass at several segment heights and both window
        widths and keep the fastest (the state does not advance); returns the
        height (get_step_pair_window gives the width)."""
        v = ctypes.c_int()
        self._check(self.lib.le2d_tune_step_pairs(self.ref, ctypes.byref(v)), "tune_step_pairs")
        return v.value

    def get_step_pairs(self):
        v = ctypes.c_int()
        self._check(self.lib.le2d_get_step_pairs(self.ref, ctypes.byref(v)), "get_step_pairs")
        return v.value

    def time_pairs(self, passes):
        """Mean ms of one two-step pass (HIP events)."""
        a = ctypes.c_double()
        self._check(self.lib.le2d_time_pairs(self.ref, passes, ctypes.byref(a)), "time_pairs")
        return a.value

    def save(self, path):
        return self.lib.le_save_task(self.ref, path.encode() if isinstance(path, str) else path)

    def dump(self, path):
        self._check(self.lib.le2d_dump(self.ref, path.encode()), "le2d_dump")

    def restore(self, path):
        self._check(self.lib.le2d_restore(self.ref, path.encode()), "le2d_restore")

    def device_bytes(self):
        return self.lib.le2d_device_bytes(self.ref)

    def set_materials(self, c1_c2_rho, ids=None):
        """c1_c2_rho: (count, 3) floats; ids: (ny, nx) uint8 or None."""
        tab = np.ascontiguousarray(c1_c2_rho, dtype=np.float64).reshape(-1, 3)
        self._tab = tab
        self._ids = None if ids is None else np.ascontiguousarray(ids, dtype=np.uint8)
        if self._ids is not None:
            assert self._ids.shape == (self.ny, self.nx)
        self._check(self.lib.le2d_set_material_plane(
            self.ref, len(tab), tab.ctypes.data,
            None if self._ids is None else self._ids.ctypes.data), "set_material_plane")

    def device_view(self):
        v = le2d_view()
        self._check(self.lib.le2d_task_view(self.ref, ctypes.byref(v)), "le2d_task_view")
        return v

    def energy(self):
        e = ctypes.c_double()
        self._check(self.lib.le2d_energy(self.ref, ctypes.byref(e)), "le2d_energy")
        return e.value

    def close(self):
        if self._open:
            self.lib.le_free_task(self.ref)
            self._open = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Batch:
    """Many independent grids advanced together (include/le2d.h, le2d_batch_*).
    State arrays are AOS numpy arrays shaped (ny, nx, 5) of the batch dtype."""

    def __init__(self, dtype=np.float32, limiter="superbee"):
        self.dtype = np.dtype(dtype)
        self.lib = load(np.float64, limiter)  # the batch API is in both libraries
        self.h = self.lib.le2d_batch_create(self.dtype.itemsize)
        if not self.h:
            raise LE2DError("le2d_batch_create: " + self.lib.le2d_last_error().decode())
        self.shapes = []

    def _check(self, rc, what):
        if rc != 0:
            raise LE2DError(f"{what}: {rc} {self.lib.le2d_last_error().decode()}")

    def add(self, nx, ny, dt, h=(1.0, 1.0), mat=(2.0, 1.0, 1.0)):
        i = self.lib.le2d_batch_add(self.h, nx, ny, dt, h[0], h[1], mat[0], mat[1], mat[2])
        if i < 0:
            self._check(-i, "le2d_batch_add")
        self.shapes.append((ny, nx))
        return i

    def commit(self):
        self._check(self.lib.le2d_batch_commit(self.h), "le2d_batch_commit")

    def device_bytes(self):
        return self.lib.le2d_batch_device_bytes(self.h)

    def set_ball(self, i, c, r, s):
        self._check(self.lib.le2d_batch_set_ball(self.h, i, c[0], c[1], r, s), "set_ball")

    def load(self, i, state):
        a = np.ascontiguousarray(state, dtype=self.dtype)
        assert a.shape == self.shapes[i] + (5,)
        self._check(self.lib.le2d_batch_load(self.h, i, a.ctypes.data), "le2d_batch_load")

    def read(self, i):
        a = np.empty(self.shapes[i] + (5,), dtype=self.dtype)
        self._check(self.lib.le2d_batch_read(self.h, i, a.ctypes.data), "le2d_batch_read")
        return a

    def step(self, n):
        self._check(self.lib.le2d_batch_step(self.h, n), "le2d_batch_step")

    def step_plain(self, n):
        self._check(self.lib.le2d_batch_step_plain(self.h, n), "le2d_batch_step_plain")

    def wait(self):
        self._check(self.lib.le2d_batch_wait(self.h), "le2d_batch_wait")

    FORMS = {"tiles": 0, "persistent": 1, "pairs": 2, "pairs_split": 3}

    def set_shape(self, fb, seg, form=None):
        """Window width and rows per segment; form ("tiles", "persistent",
        "pairs" or "pairs_split") also picks the launch form (None keeps
        it)."""
        if form is None:
            self._check(self.lib.le2d_batch_set_shape(self.h, fb, seg), "le2d_batch_set_shape")
        else:
            self._check(self.lib.le2d_batch_set_launch(self.h, fb, seg, self.FORMS[form]),
                        "le2d_batch_set_launch")

    def form(self):
        """The launch form: "tiles", "persistent", "pairs" or "pairs_split"."""
        v = ctypes.c_int()
        self._check(self.lib.le2d_batch_get_launch(self.h, ctypes.byref(v)), "get_launch")
        return {n: k for k, n in self.FORMS.items()}[v.value]

    def get_shape(self):
        v = [ctypes.c_int() for _ in range(4)]
        self._check(self.lib.le2d_batch_get_shape(self.h, *[ctypes.byref(x) for x in v]),
                    "get_shape")
        return tuple(x.value for x in v)

    def time(self, steps, plain=False):
        ms = ctypes.c_double()
        self._check(self.lib.le2d_batch_time(self.h, steps, int(plain), ctypes.byref(ms)),
                    "le2d_batch_time")
        return ms.value

    def device_view(self, i):
        v = le2d_view()
        self._check(self.lib.le2d_batch_view(self.h, i, ctypes.byref(v)), "le2d_batch_view")
        return v

    def autotune(self, trials=3):
        fb, seg, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._check(self.lib.le2d_batch_autotune(self.h, trials, ctypes.byref(fb),
                                                 ctypes.byref(seg), ctypes.byref(ms)),
                    "le2d_batch_autotune")
        return fb.value, seg.value, ms.value

    def close(self):
        if self.h:
            self.lib.le2d_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Server:
    """Serving loop (include/le2d.h, le2d_server_*): requests admitted by HBM
    budget and slots, continuous batching of their steps on the device."""

    def __init__(self, budget_bytes, max_live=256, plain=False, limiter="superbee"):
        self.lib = load(np.float64, limiter)
        self.h = self.lib.le2d_server_create(int(budget_bytes), max_live, int(plain))
        if not self.h:
            raise LE2DError("le2d_server_create: " + self.lib.le2d_last_error().decode())
        self.shapes = {}

    def _check(self, rc, what):
        if rc != 0:
            raise LE2DError(f"{what}: {rc} {self.lib.le2d_last_error().decode()}")

    def set_shape(self, fb=256, seg=16, max_chunk=16):
        self._check(self.lib.le2d_server_set_shape(self.h, fb, seg, max_chunk), "set_shape")

    def set_pairs(self, pair_seg=64):
        """Largest rows per segment of the two-step launches (the server picks
        the height for its live set); 0: one step per launch."""
        self._check(self.lib.le2d_server_set_pairs(self.h, pair_seg), "set_pairs")

    def set_pair_window(self, lanes=0):
        """Lanes per two-step workgroup: 64, 128 or 256 (0: 256 for FP32
        requests, 64 for FP64)."""
        self._check(self.lib.le2d_server_set_pair_window(self.h, lanes), "set_pair_window")

    def set_pair_split(self, on=1):
        """64-lane two-step windows: the two stages on two wavefronts."""
        self._check(self.lib.le2d_server_set_pair_split(self.h, int(on)), "set_pair_split")

    def set_admission(self, max_queue=0, window=64, max_pass=16):
        """Queue limit (0: none; submit then raises BusyError at the limit),
        the queued requests admission scans, and how often a request that
        does not fit may be passed by later ones."""
        self._check(self.lib.le2d_server_set_admission(self.h, max_queue, window, max_pass),
                    "set_admission")

    def cancel(self, i):
        self._check(self.lib.le2d_server_cancel(self.h, i), "le2d_server_cancel")

    def submit(self, nx, ny, steps, dtype=np.float32, dt=0.25, h=(1.0, 1.0),
               mat=(2.0, 1.0, 1.0), disk=(None, None, 10.0, 1.0), keep=False):
        cx, cy, r, sv = disk
        q = le2d_request(nx, ny, np.dtype(dtype).itemsize, steps, dt, h[0], h[1], mat[0],
                         mat[1], mat[2], nx / 2 if cx is None else cx,
                         ny / 2 if cy is None else cy, r, sv, int(keep))
        i = self.lib.le2d_server_submit(self.h, ctypes.byref(q))
        if i == -ERR_BUSY:
            raise BusyError("le2d_server_submit: " + self.lib.le2d_last_error().decode())
        if i < 0:
            self._check(-i, "le2d_server_submit")
        self.shapes[i] = ((ny, nx), np.dtype(dtype))
        return i

    def wait(self, i, timeout=-1.0):
        self._check(self.lib.le2d_server_wait(self.h, i, timeout), "le2d_server_wait")

    def drain(self, timeout=-1.0):
        self._check(self.lib.le2d_server_drain(self.h, timeout), "le2d_server_drain")

    def drained(self, timeout):
        """Wait up to `timeout` seconds for every submitted request: True when
        all are done, False when the time ran out (LE2D_ERR_STATE); any
        other failure raises."""
        rc = self.lib.le2d_server_drain(self.h, timeout)
        if rc == ERR_STATE:
            return False
        self._check(rc, "le2d_server_drain")
        return True

    def result(self, i):
        shape, dt = self.shapes[i]
        a = np.empty(shape + (5,), dtype=dt)
        self._check(self.lib.le2d_server_result(self.h, i, a.ctypes.data), "le2d_server_result")
        return a

    def times(self, i):
        t = [ctypes.c_double() for _ in range(3)]
        self._check(self.lib.le2d_server_times(self.h, i, *[ctypes.byref(x) for x in t]),
                    "le2d_server_times")
        return tuple(x.value for x in t)

    def stats(self):
        st = le2d_server_stats()
        self._check(self.lib.le2d_server_get_stats(self.h, ctypes.byref(st)), "stats")
        return {f: getattr(st, f) for f, _ in st._fields_}

    def close(self):
        if self.h:
            self.lib.le2d_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
