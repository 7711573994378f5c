"""ctypes mirror of the le_core.h ABI (include/le_core.h).

The structures follow the reference's simple/le_core.h:130-212 field for
field, for both builds of ``real`` (double with -DUSE_DOUBLE, float without).
Sizes on x86-64: le_task 112 B (double) / 64 B (float).
"""
import ctypes

ST_SOA = 0  # simple/le_core.h:34
ST_AOS = 1  # simple/le_core.h:35
NODE_FIELDS = ("vx", "vy", "sxx", "sxy", "syy")  # simple/le_core.h:156-159


def _make(real):
    class le_vec2(ctypes.Structure):
        _fields_ = [("x", real), ("y", real)]

    class le_point2(ctypes.Structure):
        _fields_ = [("x", ctypes.c_int), ("y", ctypes.c_int)]

    class le_material(ctypes.Structure):
        _fields_ = [(n, real) for n in
                    ("c1", "c2", "rho", "irhoc1", "irhoc2", "rhoc1", "rhoc2", "rhoc3")]

    class le_task(ctypes.Structure):
        _fields_ = [("dt", real), ("h", le_vec2), ("n", le_point2),
                    ("mat", le_material), ("stype", ctypes.c_int),
                    ("grid", ctypes.POINTER(real))]

    return le_vec2, le_point2, le_material, le_task


class ABI:
    """Structure types plus the argtypes of the le_core.h entry points."""

    def __init__(self, real):
        self.real = real
        self.is_double = real is ctypes.c_double
        (self.le_vec2, self.le_point2, self.le_material,
         self.le_task) = _make(real)

    def bind(self, lib):
        """Attach argtypes/restypes for the entry points simple/le_core.h:217-254
        declares (plus the linkable single sweeps)."""
        T, R = self.le_task, self.real
        P = ctypes.POINTER
        sig = {
            "le_init_material": ([R, R, R, P(self.le_material)], None),
            "le_init_task": ([P(T), R, self.le_vec2, self.le_material,
                              self.le_point2, ctypes.c_int], None),
            "le_free_task": ([P(T)], None),
            "le_set_ball": ([P(T), self.le_vec2, R, R], None),
            "le_save_task": ([P(T), ctypes.c_char_p], ctypes.c_int),
            "le_step": ([P(T)], None),
            "le_step_soa": ([P(T)], None),
            "le_step_x": ([P(T)], None),
            "le_step_y": ([P(T)], None),
            "le_step_y_cf": ([P(T), ctypes.c_int], None),
            "le_soa_step_x": ([P(T)], None),
            "le_soa_step_y": ([P(T)], None),
            "le_step_cf": ([P(T), ctypes.c_int], None),
        }
        for name, (args, res) in sig.items():
            fn = getattr(lib, name, None)
            if fn is None:
                continue
            fn.argtypes = args
            fn.restype = res
        return lib


ABI_D = ABI(ctypes.c_double)
ABI_F = ABI(ctypes.c_float)


def abi_for(dtype):
    """ABI for a numpy dtype / 'f8' / 'f4'."""
    import numpy as np
    return ABI_D if np.dtype(dtype) == np.float64 else ABI_F
