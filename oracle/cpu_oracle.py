"""ctypes binding of the CPU oracle (test infrastructure only).

`OracleState` owns numpy arrays for one world + particle set and binds them
to an oracle context by name. Only tests, the smoke check and the benchmark's
CPU-baseline leg import this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(bits=32):
    if bits not in _LIBS:
        path = os.path.join(HERE, "_build", "liboracle%d.so" % bits)
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        vp, d, i64, ci = ctypes.c_void_p, ctypes.c_double, ctypes.c_longlong, ctypes.c_int
        L.oc_create.restype = vp
        L.oc_create.argtypes = [ci, ci, ci, d]
        L.oc_destroy.argtypes = [vp]
        L.oc_set_ptr.argtypes = [vp, ctypes.c_char_p, vp]
        L.oc_set_f.argtypes = [vp, ctypes.c_char_p, d]
        L.oc_set_i.argtypes = [vp, ctypes.c_char_p, i64]
        L.oc_get_f.argtypes = [vp, ctypes.c_char_p]
        L.oc_get_f.restype = d
        L.oc_get_i.argtypes = [vp, ctypes.c_char_p]
        L.oc_get_i.restype = i64
        for fn in ["oc_rasterize_static", "oc_clear_faces", "oc_p2g", "oc_apply_bc", "oc_snapshot",
                   "oc_project", "oc_extrapolate", "oc_pressure_finish", "oc_pressure_solve",
                   "oc_splat_density", "oc_push_surface", "oc_vapor_collisions",
                   "oc_ice_contacts"]:
            getattr(L, fn).argtypes = [vp]
        L.oc_mark_fluid.argtypes = [vp, ci, ci]
        for fn in ["oc_viscosity", "oc_apply_gravity", "oc_vorticity", "oc_advect", "oc_substep", "oc_update_heat",
                   "oc_update_evaporation", "oc_update_currents", "oc_update_phase"]:
            getattr(L, fn).argtypes = [vp, d]
        L.oc_g2p.argtypes = [vp, d]
        L.oc_g2p.restype = ci
        L.oc_pressure_assemble.argtypes = [vp]
        L.oc_pressure_assemble.restype = ci
        L.oc_sor_lex.argtypes = [vp, ci]
        L.oc_sor_redblack.argtypes = [vp, ci]
        L.oc_residual.argtypes = [vp, ctypes.POINTER(d)]
        L.oc_residual.restype = d
        L.oc_clamp_umax.argtypes = [vp]
        L.oc_clamp_umax.restype = d
        L.oc_dt_law.argtypes = [vp]
        L.oc_dt_law.restype = d
        L.oc_sample_face.argtypes = [vp, vp, ci, d, d, d]
        L.oc_sample_face.restype = d
        L.oc_sample_vel3.argtypes = [vp, d, d, d, ctypes.POINTER(d)]
        L.oc_terrain_radius_at.argtypes = [vp, d, d, d]
        L.oc_terrain_radius_at.restype = d
        L.oc_rock_cell_at.argtypes = [vp, d, d, d]
        L.oc_rock_cell_at.restype = ci
        L.oc_mt_build.argtypes = [vp, ci, ci, ci, d, d, vp, vp, ctypes.c_long, ctypes.POINTER(d)]
        L.oc_mt_build.restype = ctypes.c_long
        L.oc_autotune_iso.argtypes = [vp]
        L.oc_autotune_iso.restype = d
        L.oc_kinetic_energy.argtypes = [vp]
        L.oc_kinetic_energy.restype = d
        L.oc_set_balls.argtypes = [vp, ci, vp]
        L.oc_set_balls.restype = ci
        L.oc_get_balls.argtypes = [vp, ci, vp]
        L.oc_get_balls.restype = ci
        L.oc_push_balls.argtypes = [vp]
        L.oc_update_balls.argtypes = [vp, d]
        L.oc_ball_hist.argtypes = [vp]
        L.oc_ball_hist.restype = ctypes.POINTER(ctypes.c_float)
        L.oc_metaball_params.argtypes = [ci, d, d, d, d, ctypes.POINTER(ci), ctypes.POINTER(d), ctypes.POINTER(d),
                                         ctypes.POINTER(ci), ctypes.POINTER(d)]
        L.oc_metaball_field.argtypes = [vp, vp, vp, vp, ci, ci, d, d, ci, vp]
        L.oc_ball_surf.argtypes = [vp]
        L.oc_ball_surf.restype = ctypes.POINTER(ctypes.c_float)
        _LIBS[bits] = L
    return _LIBS[bits]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


GRID_F = ["u", "v", "w", "uW", "vW", "wW", "uO", "vO", "wO"]
PART_F = ["px", "py", "pz", "pvx", "pvy", "pvz", "pT", "pAir", "pWx", "pWy", "pWz", "pLight", "pPh"]


class OracleState:
    """World + particle state in numpy, bound to one oracle context."""

    def __init__(self, world, bits=32):
        self.L = lib(bits)
        self.bits = bits
        self.real = np.float32 if bits == 32 else np.float64
        nx, ny, nz, dx = world["nx"], world["ny"], world["nz"], world["dx"]
        self.nx, self.ny, self.nz, self.dx = nx, ny, nz, dx
        self.ctx = ctypes.c_void_p(self.L.oc_create(nx, ny, nz, dx))
        self.nC = nx * ny * nz
        self.uN, self.vN, self.wN = (nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)
        self.nD = (nx + 1) * (ny + 1) * (nz + 1)
        self.arrays = {}
        sizes = {"u": self.uN, "v": self.vN, "w": self.wN}
        for k in GRID_F:
            self._bind(k, np.zeros(sizes[k[0]], self.real))
        self._bind("validU", np.zeros(self.uN, np.uint8))
        self._bind("validV", np.zeros(self.vN, np.uint8))
        self._bind("validW", np.zeros(self.wN, np.uint8))
        self._bind("cellType", np.zeros(self.nC, np.uint8))
        for k in ["q", "rhs", "diag", "cellPhi"]:
            self._bind(k, np.zeros(self.nC, self.real))
        self._bind("dens", np.zeros(self.nD, self.real))
        for k in ["coreR", "oceanR", "atmosphereH", "spacing", "particleVolume", "dv"]:
            if k in world:
                self.set_f(k, world[k])
        if "tn" in world:
            self.set_i("tn", world["tn"])

    def __del__(self):
        try:
            self.L.oc_destroy(self.ctx)
        except Exception:
            pass

    def _bind(self, name, arr):
        arr = np.ascontiguousarray(arr)
        self.arrays[name] = arr
        if self.L.oc_set_ptr(self.ctx, name.encode(), _ptr(arr)) != 0:
            raise KeyError(name)
        return arr

    def bind(self, name, arr, dtype=None):
        if dtype is not None:
            arr = np.asarray(arr, dtype=dtype)
        return self._bind(name, np.array(arr, copy=True))

    def __getitem__(self, k):
        return self.arrays[k]

    def set_f(self, k, v):
        if self.L.oc_set_f(self.ctx, k.encode(), float(v)) != 0:
            raise KeyError(k)

    def set_i(self, k, v):
        if self.L.oc_set_i(self.ctx, k.encode(), int(v)) != 0:
            raise KeyError(k)

    def get_f(self, k):
        return self.L.oc_get_f(self.ctx, k.encode())

    def get_i(self, k):
        return self.L.oc_get_i(self.ctx, k.encode())

    def set_static(self, terrain_solid, terrain_R, static_types, gu, gv, gw):
        self.bind("tsolid", terrain_solid, np.uint8)
        self.bind("tR", terrain_R, np.float32)
        if static_types is not None:
            self.bind("staticTypes", static_types, np.uint8)
        self.bind("gu", gu, self.real)
        self.bind("gv", gv, self.real)
        self.bind("gw", gw, self.real)

    def set_particles(self, n, cap=None, **arrs):
        cap = max(cap or n, n)
        self.set_i("nP", n)
        for k in PART_F:
            a = np.zeros(cap, self.real)
            if k in arrs:
                a[:n] = arrs[k][:n]
            self._bind(k, a)
        for k in ["pflag", "pcool", "pDepth"]:
            a = np.zeros(cap, np.uint8)
            if k in arrs:
                a[:n] = arrs[k][:n]
            self._bind(k, a)
        for k in ["advU", "advV", "advW"]:
            self._bind(k, np.zeros(cap, np.float64))

    def set_params(self, **kw):
        ints = {"pressureIters", "ceilReflect", "hasSun", "cfS", "bfS", "evS", "nP", "tn", "sorRedBlack",
                "residualEvery", "maxPressureIters", "spinOn"}
        for k, v in kw.items():
            (self.set_i if k in ints else self.set_f)(k, v)

    def call(self, fn, *args):
        return getattr(self.L, fn)(self.ctx, *args)

    def set_balls(self, balls):
        """balls: (n, 9) doubles x y z vx vy vz r rho submerged; binds the
        solid-velocity fields on first use"""
        b = np.ascontiguousarray(np.asarray(balls, np.float64).reshape(-1, 9))
        if "svx" not in self.arrays:
            for k in ["svx", "svy", "svz"]:
                self._bind(k, np.zeros(self.nC, self.real))
        if self.L.oc_set_balls(self.ctx, len(b), _ptr(b)) != 0:
            raise ValueError("too many balls")

    def balls(self):
        n = self.get_i("nBalls")
        out = np.zeros((n, 9), np.float64)
        self.L.oc_get_balls(self.ctx, n, _ptr(out))
        return out

    def ball_hist(self):
        n = self.get_i("nBalls")
        h = np.ctypeslib.as_array(self.L.oc_ball_hist(self.ctx), shape=(max(n, 1) * 24,))[:n * 24].copy()
        sy = np.ctypeslib.as_array(self.L.oc_ball_surf(self.ctx), shape=(max(n, 1),))[:n].copy()
        return h.reshape(n, 24), sy


def mt_build(dens, n, dx, iso, cap=None, bits=32):
    """Marching tetrahedra on a corner field of (n+1)^3; returns count, volume, pos, nrm."""
    L = lib(bits)
    real = np.float32 if bits == 32 else np.float64
    nx, ny, nz = n if isinstance(n, tuple) else (n, n, n)
    dens = np.ascontiguousarray(dens, dtype=real)
    vol = ctypes.c_double(0)
    cnt = L.oc_mt_build(_ptr(dens), nx, ny, nz, dx, iso, None, None, 0, ctypes.byref(vol))
    cap = cnt if cap is None else cap
    pos = np.zeros(max(cap, 1) * 3, np.float32)
    nrm = np.zeros(max(cap, 1) * 3, np.float32)
    cnt = L.oc_mt_build(_ptr(dens), nx, ny, nz, dx, iso, _ptr(pos), _ptr(nrm), cap, ctypes.byref(vol))
    return cnt, vol.value, pos[:cnt * 3], nrm[:cnt * 3]


def metaball_params(nx, dx, spacing, tension, detail=float("nan")):
    """(m, mdx, R, rad, iso) of Metaballs.build (surface.js:178-250)."""
    L = lib(32)
    m, rad = ctypes.c_int(), ctypes.c_int()
    mdx, R, iso = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    L.oc_metaball_params(nx, dx, spacing, tension, detail, ctypes.byref(m), ctypes.byref(mdx), ctypes.byref(R),
                         ctypes.byref(rad), ctypes.byref(iso))
    return m.value, mdx.value, R.value, rad.value, iso.value


def metaball_field(px, py, pz, pflag, m, mdx, R, rad, bits=32):
    """The metaball corner field, (m+1)^3 values."""
    L = lib(bits)
    real = np.float32 if bits == 32 else np.float64
    px, py, pz = (np.ascontiguousarray(a, dtype=real) for a in (px, py, pz))
    fl = np.ascontiguousarray(pflag, dtype=np.uint8)
    field = np.zeros((m + 1) ** 3, real)
    L.oc_metaball_field(_ptr(px), _ptr(py), _ptr(pz), _ptr(fl), len(px), m, mdx, R, rad, _ptr(field))
    return field
