"""ctypes binding of the watersim C ABI (include/watersim.h).

This is the Python host for tests and the benchmark; the JavaScript host
(js/solver.js with the N-API addon) binds the same entry points. There is no
CPU fallback: if the HIP library is missing or no GPU is present, creating a
Solver raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libwatersim_hip.so")

STAGES = {
    "clear": 0, "rasterize": 1, "p2g": 2, "bc": 3, "snapshot": 4, "viscosity": 5, "gravity": 6,
    "pressure": 7, "project": 8, "vorticity": 9, "clamp": 10, "extrapolate": 11, "g2p": 12, "advect": 13,
}

# exported entry points, as declared in include/watersim.h
EXPORTS = ["ws_create", "ws_destroy", "ws_last_error", "ws_set_f", "ws_get_f", "ws_upload", "ws_download",
           "ws_substep", "ws_stage", "ws_step_frame", "ws_run_substeps", "ws_splat_density", "ws_mt_build",
           "ws_autotune_iso", "ws_kinetic_energy", "ws_sync"]

_lib = None


class WatersimError(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Loads the library (no device call). Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise WatersimError("HIP library not built: %s (run __graft_entry__.build())" % path)
    L = ctypes.CDLL(path)
    vp, d, ci, cs, sz = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t
    L.ws_create.argtypes = [ctypes.POINTER(vp), ci, ci, ci, d, ci]
    L.ws_destroy.argtypes = [vp]
    L.ws_destroy.restype = None
    L.ws_last_error.argtypes = [vp]
    L.ws_last_error.restype = cs
    L.ws_set_f.argtypes = [vp, cs, d]
    L.ws_get_f.argtypes = [vp, cs, ctypes.POINTER(d)]
    L.ws_upload.argtypes = [vp, cs, vp, sz]
    L.ws_download.argtypes = [vp, cs, vp, sz]
    L.ws_substep.argtypes = [vp, d]
    L.ws_stage.argtypes = [vp, ci, d]
    L.ws_step_frame.argtypes = [vp, d, ctypes.POINTER(ci)]
    L.ws_run_substeps.argtypes = [vp, ci]
    L.ws_splat_density.argtypes = [vp]
    L.ws_mt_build.argtypes = [vp, vp, ci, ci, ci, d, d, vp, vp, ctypes.c_long, ctypes.POINTER(ctypes.c_long),
                              ctypes.POINTER(d)]
    L.ws_autotune_iso.argtypes = [vp, ctypes.POINTER(d)]
    L.ws_kinetic_energy.argtypes = [vp, ctypes.POINTER(d)]
    L.ws_sync.argtypes = [vp]
    for f in EXPORTS:
        if f not in ("ws_destroy", "ws_last_error"):
            getattr(L, f).restype = ci
    _lib = L
    return L


PART_F32 = ["px", "py", "pz", "pvx", "pvy", "pvz", "pT", "pAir", "advU", "advV", "advW"]
PART_U8 = ["pflag", "pcool"]
CELL_U8 = ["cellType", "staticTypes"]
CELL_F32 = ["q", "rhs"]
FACE_F32 = ["u", "v", "w", "uO", "vO", "wO", "uW", "vW", "wW", "gu", "gv", "gw"]
FACE_U8 = ["validU", "validV", "validW"]


class Solver:
    """One simulation context on the GPU (FluidSolver's device state)."""

    def __init__(self, nx, ny, nz, dx, capacity):
        self.L = load()
        self.ctx = ctypes.c_void_p()
        rc = self.L.ws_create(ctypes.byref(self.ctx), nx, ny, nz, dx, capacity)
        if rc != 0:
            raise WatersimError("ws_create failed (%d): no usable GPU or allocation failure" % rc)
        self.nx, self.ny, self.nz, self.dx, self.capacity = nx, ny, nz, dx, capacity
        self.nC = nx * ny * nz
        self.sizes = {"u": (nx + 1) * ny * nz, "v": nx * (ny + 1) * nz, "w": nx * ny * (nz + 1)}
        self.nD = (nx + 1) * (ny + 1) * (nz + 1)
        self.nP = 0

    def close(self):
        if self.ctx:
            self.L.ws_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = self.L.ws_last_error(self.ctx)
            raise WatersimError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def set(self, **kw):
        for k, v in kw.items():
            self._check(self.L.ws_set_f(self.ctx, k.encode(), float(v)), "set " + k)
            if k == "nP":
                self.nP = int(v)

    def get(self, k):
        out = ctypes.c_double()
        self._check(self.L.ws_get_f(self.ctx, k.encode(), ctypes.byref(out)), "get " + k)
        return out.value

    def _shape(self, name):
        if name in PART_F32:
            return self.nP, np.float32
        if name in PART_U8:
            return self.nP, np.uint8
        if name in CELL_U8:
            return self.nC, np.uint8
        if name in CELL_F32:
            return self.nC, np.float32
        if name in FACE_F32:
            comp = name[1] if name[0] == "g" else name[0]
            return self.sizes[comp], np.float32
        if name in FACE_U8:
            return self.sizes[name[-1].lower()], np.uint8
        if name == "dens":
            return self.nD, np.float32
        if name == "terrainR":
            tn = int(self.get("tn"))
            return tn ** 3, np.float32
        raise KeyError(name)

    def upload(self, name, arr):
        n, dt = self._shape(name)
        a = np.ascontiguousarray(arr, dtype=dt)
        if a.size != n:
            raise ValueError("%s: expected %d elements, got %d" % (name, n, a.size))
        self._check(self.L.ws_upload(self.ctx, name.encode(), a.ctypes.data_as(ctypes.c_void_p), a.nbytes),
                    "upload " + name)

    def download(self, name):
        n, dt = self._shape(name)
        a = np.empty(n, dt)
        self._check(self.L.ws_download(self.ctx, name.encode(), a.ctypes.data_as(ctypes.c_void_p), a.nbytes),
                    "download " + name)
        return a

    def set_rng(self, cf, bf, ev):
        a = np.array([cf, bf, ev], np.uint32)
        self._check(self.L.ws_upload(self.ctx, b"rng", a.ctypes.data_as(ctypes.c_void_p), 12), "upload rng")

    def get_rng(self):
        a = np.zeros(3, np.uint32)
        self._check(self.L.ws_download(self.ctx, b"rng", a.ctypes.data_as(ctypes.c_void_p), 12), "download rng")
        return a

    def substep(self, dt):
        self._check(self.L.ws_substep(self.ctx, dt), "substep")

    def stage(self, name, dt=0.0):
        self._check(self.L.ws_stage(self.ctx, STAGES[name], dt), "stage " + name)

    def step(self, dt_frame):
        n = ctypes.c_int()
        self._check(self.L.ws_step_frame(self.ctx, dt_frame, ctypes.byref(n)), "step")
        return n.value

    def run_substeps(self, n):
        self._check(self.L.ws_run_substeps(self.ctx, n), "run_substeps")

    def sync(self):
        self._check(self.L.ws_sync(self.ctx), "sync")

    def splat_density(self):
        self._check(self.L.ws_splat_density(self.ctx), "splat_density")

    def mt_build(self, dens=None, dims=None, dx=None, iso=1.8, want_mesh=True):
        nx, ny, nz = dims if dims else (self.nx, self.ny, self.nz)
        dx = self.dx if dx is None else dx
        cnt = ctypes.c_long()
        vol = ctypes.c_double()
        dptr = None
        if dens is not None:
            dens = np.ascontiguousarray(dens, np.float32)
            dptr = dens.ctypes.data_as(ctypes.c_void_p)
        self._check(self.L.ws_mt_build(self.ctx, dptr, nx, ny, nz, dx, iso, None, None, 0, ctypes.byref(cnt),
                                       ctypes.byref(vol)), "mt_build")
        if not want_mesh:
            return cnt.value, vol.value, None, None
        n = cnt.value
        pos = np.zeros(max(n, 1) * 3, np.float32)
        nrm = np.zeros(max(n, 1) * 3, np.float32)
        self._check(self.L.ws_mt_build(self.ctx, dptr, nx, ny, nz, dx, iso, pos.ctypes.data_as(ctypes.c_void_p),
                                       nrm.ctypes.data_as(ctypes.c_void_p), n, ctypes.byref(cnt),
                                       ctypes.byref(vol)), "mt_build")
        return cnt.value, vol.value, pos[:n * 3], nrm[:n * 3]

    def autotune_iso(self):
        out = ctypes.c_double()
        self._check(self.L.ws_autotune_iso(self.ctx, ctypes.byref(out)), "autotune_iso")
        return out.value

    def kinetic_energy(self):
        out = ctypes.c_double()
        self._check(self.L.ws_kinetic_energy(self.ctx, ctypes.byref(out)), "kinetic_energy")
        return out.value
