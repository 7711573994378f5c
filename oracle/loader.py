"""ctypes loaders for the oracle and for the compiled reference (test infrastructure only)."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EDGE_CLAMP = 0
EDGE_RING = 1


def build_oracle(ref=False):
    """Compile the oracle (and, if asked and the sources exist, the reference)."""
    subprocess.check_call(["make", "-s", "-B", "-C", HERE, "all"])
    if ref and os.path.isdir("/root/reference/simple"):
        subprocess.check_call(["make", "-s", "-B", "-C", HERE, "ref"])


def _ptr(a, real):
    return a.ctypes.data_as(ctypes.POINTER(real))


class Oracle:
    """Numpy front end of le2d_ref.c.  Grids are numpy arrays shaped (ny, nx, 5)
    (AOS, node (i, j) at [j, i]) or (5, ny, nx) (SOA planes)."""

    LIMITERS = {"superbee": 0, "minmod": 1, "cir": 2}

    def __init__(self, dtype=np.float64, limiter="superbee"):
        """`limiter`: the build's flux limiter (the reference chooses it at
        compile time, simple/le_core.c:36-43): superbee, minmod or cir."""
        self.dtype = np.dtype(dtype)
        self.real = ctypes.c_double if self.dtype == np.float64 else ctypes.c_float
        if limiter not in self.LIMITERS:
            raise ValueError(f"limiter {limiter!r}: superbee, minmod or cir")
        self.limiter_name = limiter
        stem = "liboracle" if limiter == "superbee" else "liboracle_" + limiter
        name = stem + ("_d.so" if self.dtype == np.float64 else "_f.so")
        # LE2D_ORACLE_LIBDIR: another build of the oracle (oracle/_san, the
        # sanitizer build of `make san`)
        path = os.path.join(os.environ.get("LE2D_ORACLE_LIBDIR") or HERE, name)
        if not os.path.exists(path):
            build_oracle()
        self.lib = ctypes.CDLL(path)
        R, P = self.real, ctypes.POINTER
        self.Mat = type("orc_material", (ctypes.Structure,),
                        {"_fields_": [(n, R) for n in ("c1", "c2", "rho", "irhoc1", "irhoc2",
                                                        "rhoc1", "rhoc2", "rhoc3")]})
        L = self.lib
        L.orc_material_init.argtypes = [R, R, R, P(self.Mat)]
        L.orc_superbee.argtypes = [R]
        L.orc_superbee.restype = R
        L.orc_limiter.argtypes = [R]
        L.orc_limiter.restype = R
        if L.orc_limiter_kind() != self.LIMITERS[limiter]:
            raise RuntimeError(f"{path} was not built with the {limiter} limiter")
        L.orc_tvd2.argtypes = [R] * 5
        L.orc_tvd2.restype = R
        grid_args = [P(R), ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_sweep_x.argtypes = grid_args + [R, R, P(self.Mat)]
        L.orc_sweep_y.argtypes = grid_args + [R, R, P(self.Mat), ctypes.c_int]
        L.orc_step.argtypes = grid_args + [R, R, R, P(self.Mat)]
        L.orc_run.argtypes = grid_args + [R, R, R, P(self.Mat), ctypes.c_int]
        L.orc_set_ball.argtypes = grid_args + [R] * 6
        L.orc_set_ball.restype = ctypes.c_long
        L.orc_save_vtk.argtypes = [P(R), ctypes.c_int, ctypes.c_int, ctypes.c_int, R, R,
                                   ctypes.c_char_p]
        L.orc_step_mat.argtypes = grid_args + [R, R, R, ctypes.c_int, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_int]
        L.orc_sweep_mat.argtypes = grid_args + [R, R, ctypes.c_int, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        assert L.orc_real_size() == self.dtype.itemsize

    # --- scalar pieces -----------------------------------------------------
    def material(self, c1, c2, rho):
        m = self.Mat()
        self.lib.orc_material_init(c1, c2, rho, ctypes.byref(m))
        return m

    def superbee(self, r):
        return self.lib.orc_superbee(r)

    def limiter(self, r):
        """The build's limiter (superbee, minmod or CIR) at r."""
        return self.lib.orc_limiter(r)

    def tvd2(self, c, a, b, u, d):
        return self.lib.orc_tvd2(c, a, b, u, d)

    # --- grid operations (in place) ----------------------------------------
    def _grid(self, g, soa=None):
        """(nx, ny, soa) of an AOS (ny, nx, 5) or SOA (5, ny, nx) grid.  A
        (5, n, 5) array is either an AOS grid 5 rows high or an SOA grid 5
        nodes wide; without `soa` it is taken as AOS, so SOA grids of width 5
        need soa=True."""
        assert g.dtype == self.dtype and g.flags.c_contiguous and g.ndim == 3
        if soa is None:
            soa = g.shape[-1] != 5
        if soa:
            assert g.shape[0] == 5, g.shape
            return g.shape[2], g.shape[1], 1
        assert g.shape[-1] == 5, g.shape
        return g.shape[1], g.shape[0], 0

    def _mat(self, mat):
        return mat if isinstance(mat, self.Mat) else self.material(*mat)

    def sweep_x(self, g, dt, hx, mat, soa=None):
        nx, ny, soa = self._grid(g, soa)
        rc = self.lib.orc_sweep_x(_ptr(g, self.real), nx, ny, soa, dt, hx,
                                  ctypes.byref(self._mat(mat)))
        assert rc == 0

    def sweep_y(self, g, dt, hy, mat, edge=EDGE_CLAMP, soa=None):
        nx, ny, soa = self._grid(g, soa)
        rc = self.lib.orc_sweep_y(_ptr(g, self.real), nx, ny, soa, dt, hy,
                                  ctypes.byref(self._mat(mat)), edge)
        assert rc == 0

    def run(self, g, dt, h, mat, steps, soa=None):
        nx, ny, soa = self._grid(g, soa)
        rc = self.lib.orc_run(_ptr(g, self.real), nx, ny, soa, dt, h[0], h[1],
                              ctypes.byref(self._mat(mat)), steps)
        assert rc == 0

    @staticmethod
    def _table(c1c2rho, ids, nx, ny):
        tab = np.ascontiguousarray(c1c2rho, dtype=np.float64).reshape(-1, 3)
        ids = np.ascontiguousarray(ids, dtype=np.uint8)
        assert ids.shape == (ny, nx) and int(ids.max()) < len(tab)
        return tab, ids

    def run_mat(self, g, dt, h, c1c2rho, ids, steps, edge=EDGE_CLAMP, soa=None):
        """Heterogeneous scheme: ids (ny, nx) uint8 material indices into the
        (count, 3) table of (c1, c2, rho)."""
        nx, ny, soa = self._grid(g, soa)
        tab, ids = self._table(c1c2rho, ids, nx, ny)
        for _ in range(steps):
            rc = self.lib.orc_step_mat(_ptr(g, self.real), nx, ny, soa, dt, h[0], h[1], len(tab),
                                       tab.ctypes.data, ids.ctypes.data, edge)
            assert rc == 0

    def sweep_mat(self, g, dt, h, c1c2rho, ids, axis, edge=EDGE_CLAMP, soa=None):
        """One heterogeneous sweep along x (axis 0, spacing h) or y (axis 1)."""
        nx, ny, soa = self._grid(g, soa)
        tab, ids = self._table(c1c2rho, ids, nx, ny)
        rc = self.lib.orc_sweep_mat(_ptr(g, self.real), nx, ny, soa, dt, h, len(tab),
                                    tab.ctypes.data, ids.ctypes.data, axis, edge)
        assert rc == 0

    @staticmethod
    def energy_density(v, rho, c1, c2):
        """FP64 elastic energy per unit area of nodes v (..., 5) in plane strain:
        rho/2 |v|^2 + (sxx+syy)^2 / (8 (lam+mu)) + ((sxx-syy)^2/2 + 2 sxy^2) / (4 mu),
        mu = rho c2^2, lam = rho c1^2 - 2 mu (rho, c1, c2 scalars or arrays
        broadcast against v[..., 0])."""
        v = np.asarray(v, dtype=np.float64)
        mu = rho * c2 * c2
        lam = rho * c1 * c1 - 2.0 * mu
        tr = v[..., 2] + v[..., 4]
        dv = v[..., 2] - v[..., 4]
        return (0.5 * rho * (v[..., 0] ** 2 + v[..., 1] ** 2) + tr * tr / (8.0 * (lam + mu))
                + (0.5 * dv * dv + 2.0 * v[..., 3] ** 2) / (4.0 * mu))

    def energy(self, g, h, mat):
        """FP64 elastic energy of an AOS grid (homogeneous material)."""
        m = self._mat(mat)
        dens = self.energy_density(g.reshape(-1, 5), float(m.rho), float(m.c1), float(m.c2))
        return float(dens.sum()) * h[0] * h[1]

    def energy_mat(self, g, h, c1c2rho, ids):
        """FP64 elastic energy of an AOS grid with one material index per node."""
        tab = np.asarray(c1c2rho, dtype=np.float64).reshape(-1, 3)[np.asarray(ids).reshape(-1)]
        dens = self.energy_density(g.reshape(-1, 5), tab[:, 2], tab[:, 0], tab[:, 1])
        return float(dens.sum()) * h[0] * h[1]

    def set_ball(self, g, h, c, r, s, soa=None):
        nx, ny, soa = self._grid(g, soa)
        return self.lib.orc_set_ball(_ptr(g, self.real), nx, ny, soa, h[0], h[1], c[0], c[1],
                                     r, s)

    def save_vtk(self, g, h, path, soa=None):
        nx, ny, soa = self._grid(g, soa)
        return self.lib.orc_save_vtk(_ptr(g, self.real), nx, ny, soa, h[0], h[1],
                                     path.encode() if isinstance(path, str) else path)


class Reference:
    """The reference's simple/ path (oracle/_ref/libsimple_{d,f}.so), driven
    through its own le_core.h entry points with a ctypes le_task."""

    def __init__(self, dtype=np.float64):
        import sys
        sys.path.insert(0, os.path.dirname(HERE))
        from le2d_amd.abi import ABI_D, ABI_F
        self.dtype = np.dtype(dtype)
        self.abi = ABI_D if self.dtype == np.float64 else ABI_F
        name = "libsimple_d.so" if self.dtype == np.float64 else "libsimple_f.so"
        path = os.path.join(HERE, "_ref", name)
        if not os.path.exists(path):
            raise FileNotFoundError(path + " (run `make -C oracle ref` where the reference "
                                    "sources are present)")
        self.lib = self.abi.bind(ctypes.CDLL(path))

    @staticmethod
    def available(dtype=np.float64):
        name = "libsimple_d.so" if np.dtype(dtype) == np.float64 else "libsimple_f.so"
        return os.path.exists(os.path.join(HERE, "_ref", name))

    def material(self, c1, c2, rho):
        m = self.abi.le_material()
        self.lib.le_init_material(c1, c2, rho, ctypes.byref(m))
        return m

    def task(self, nx, ny, dt, h, mat, stype=1):
        t = self.abi.le_task()
        if not isinstance(mat, self.abi.le_material):
            mat = self.material(*mat)
        self.lib.le_init_task(ctypes.byref(t), dt, self.abi.le_vec2(*h), mat,
                              self.abi.le_point2(nx, ny), stype)
        return t

    def view(self, t):
        """Numpy view of the task grid: (ny, nx, 5) for AOS, (5, ny, nx) for SOA."""
        nx, ny = t.n.x, t.n.y
        flat = np.ctypeslib.as_array(t.grid, shape=(nx * ny * 5,))
        return flat.reshape(ny, nx, 5) if t.stype == 1 else flat.reshape(5, ny, nx)
