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
    "heat": 14, "evaporation": 15, "currents": 16, "phase": 17, "push_surface": 18, "vapor_collisions": 19,
    "ice_contacts": 20, "push_balls": 21, "update_balls": 22,
}

# exported entry points, as declared in include/watersim.h, watersim_world.h
# and watersim_learn.h
EXPORTS = ["ws_create", "ws_destroy", "ws_last_error", "ws_set_f", "ws_get_f", "ws_upload", "ws_download",
           "ws_substep", "ws_stage", "ws_step_frame", "ws_run_substeps", "ws_splat_density", "ws_mt_build",
           "ws_autotune_iso", "ws_kinetic_energy", "ws_sync", "ws_world_build", "ws_world_free",
           "ws_net_param_count", "ws_net_set_params", "ws_net_get_params", "ws_net_forward", "ws_net_grad",
           "ws_net_bank", "ws_net_bank_count", "ws_net_train", "ws_set_balls", "ws_get_balls", "ws_metaballs_build", "ws_read_render", "ws_tune",
           "ws_batch_create", "ws_batch_run", "ws_batch_set", "ws_batch_get", "ws_batch_last_error",
           "ws_batch_destroy", "ws_mt_build_noalloc", "ws_mark_fluid", "ws_sample_vel", "ws_planet_gravity"]


class RenderView(ctypes.Structure):
    _fields_ = [("nP", ctypes.c_int)] + [(k, ctypes.POINTER(ctypes.c_float)) for k in
                                          ("px", "py", "pz", "pT", "pLight", "pPh", "pWx", "pWy", "pWz")] + \
               [(k, ctypes.POINTER(ctypes.c_uint8)) for k in ("pflag", "pDepth")]


RENDER_BITS = {"px": 1, "py": 2, "pz": 4, "pflag": 8, "pT": 16, "pLight": 32, "pDepth": 64, "pPh": 128, "pWx": 256,
               "pWy": 512, "pWz": 1024}


class WorldParams(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("dx", ctypes.c_double),
                ("coreR", ctypes.c_double), ("depth", ctypes.c_double), ("bumpiness", ctypes.c_double),
                ("seed", ctypes.c_uint32), ("targetParticles", ctypes.c_int)]


class WorldC(ctypes.Structure):
    _fields_ = [("tn", ctypes.c_int), ("dv", ctypes.c_double), ("solid", ctypes.POINTER(ctypes.c_uint8)),
                ("R", ctypes.POINTER(ctypes.c_float)), ("staticTypes", ctypes.POINTER(ctypes.c_uint8)),
                ("gravity", ctypes.POINTER(ctypes.c_float)), ("nP", ctypes.c_long), ("capacity", ctypes.c_long),
                ("px", ctypes.POINTER(ctypes.c_float)), ("py", ctypes.POINTER(ctypes.c_float)),
                ("pz", ctypes.POINTER(ctypes.c_float)), ("spacing", ctypes.c_double),
                ("particleVolume", ctypes.c_double), ("oceanR", ctypes.c_double), ("coreR", ctypes.c_double),
                ("Rlo", ctypes.c_double), ("Rhi", ctypes.c_double), ("landFrac", ctypes.c_double)]

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
    L.ws_mt_build_noalloc.argtypes = [vp, vp, ci, ci, ci, d, d, vp, vp, ctypes.c_long, ctypes.POINTER(ctypes.c_long),
                                      ctypes.POINTER(d)]
    L.ws_mark_fluid.argtypes = [vp, ci, ci]
    L.ws_sample_vel.argtypes = [vp, vp, ctypes.c_long, vp]
    L.ws_planet_gravity.argtypes = [ci, ci, ci, d, d, vp]
    L.ws_autotune_iso.argtypes = [vp, ctypes.POINTER(d)]
    L.ws_kinetic_energy.argtypes = [vp, ctypes.POINTER(d)]
    L.ws_sync.argtypes = [vp]
    L.ws_world_build.argtypes = [ctypes.POINTER(WorldParams), ctypes.POINTER(WorldC)]
    L.ws_world_free.argtypes = [ctypes.POINTER(WorldC)]
    L.ws_world_free.restype = None
    fp = ctypes.POINTER(ctypes.c_float)
    L.ws_net_param_count.argtypes = []
    L.ws_net_set_params.argtypes = [vp, fp, ctypes.c_long]
    L.ws_net_get_params.argtypes = [vp, fp, ctypes.c_long]
    L.ws_net_forward.argtypes = [vp, ci, ci, ci, fp, fp]
    L.ws_net_grad.argtypes = [vp, ci, ci, ci, fp, fp, ctypes.POINTER(ctypes.c_uint8), fp, ctypes.POINTER(d)]
    L.ws_net_bank.argtypes = [vp, ci]
    L.ws_net_bank_count.argtypes = [vp, ctypes.POINTER(ci)]
    L.ws_net_train.argtypes = [vp, ci, ci, ci, d, ctypes.c_uint, ctypes.POINTER(d)]
    L.ws_metaballs_build.argtypes = [vp, d, d, vp, vp, ctypes.c_long, ctypes.POINTER(ctypes.c_long),
                                     ctypes.POINTER(d), ctypes.POINTER(d), ctypes.POINTER(ci)]
    L.ws_read_render.argtypes = [vp, ctypes.c_uint, ctypes.POINTER(RenderView)]
    L.ws_tune.argtypes = [vp, ci, ctypes.c_char_p, sz]
    L.ws_set_balls.argtypes = [vp, ci, vp]
    L.ws_get_balls.argtypes = [vp, ci, vp]
    L.ws_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(vp), ci, ci, ci]
    L.ws_batch_run.argtypes = [vp, ci]
    L.ws_batch_set.argtypes = [vp, cs, d]
    L.ws_batch_get.argtypes = [vp, cs, ctypes.POINTER(d)]
    L.ws_batch_last_error.argtypes = [vp]
    L.ws_batch_last_error.restype = cs
    L.ws_batch_destroy.argtypes = [vp]
    L.ws_batch_destroy.restype = None
    for f in EXPORTS:
        if f not in ("ws_destroy", "ws_last_error", "ws_world_free", "ws_batch_last_error", "ws_batch_destroy"):
            getattr(L, f).restype = ci
    L.ws_net_param_count.restype = ctypes.c_long
    _lib = L
    return L


PART_F32 = ["px", "py", "pz", "pvx", "pvy", "pvz", "pT", "pAir", "advU", "advV", "advW", "pLight", "pWx", "pWy",
            "pWz", "pPh"]
PART_U8 = ["pflag", "pcool", "pDepth"]
CELL_U8 = ["cellType", "staticTypes"]
CELL_F32 = ["q", "rhs", "svx", "svy", "svz"]
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
        self._mt_pos = self._mt_nrm = None

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
        if name == "pcell":
            return self.nP, np.int32
        if name == "terrainR":
            tn = int(self.get("tn"))
            return tn ** 3, np.float32
        if name == "terrainSolid":
            tn = int(self.get("tn"))
            return tn ** 3, np.uint8
        if name == "ballHist":
            return 25 * int(self.get("nBalls")), np.uint32
        if name == "ballSurf":
            return int(self.get("nBalls")), np.float32
        if name == "metaballField":
            return (int(self.get("mbM")) + 1) ** 3, np.float32
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

    def tune(self, sweeps=64):
        """The autotuner: fastest SOR kernel (and convolution tiling with the
        warm start on) on the current state; returns its report."""
        buf = ctypes.create_string_buffer(1024)
        self._check(self.L.ws_tune(self.ctx, sweeps, buf, len(buf)), "tune")
        return buf.value.decode()

    def read_render(self, names=tuple(RENDER_BITS)):
        """The render fields in one batched readback; returns numpy views of
        the context's pinned host memory (valid until the next call)."""
        mask = 0
        for k in names:
            mask |= RENDER_BITS[k]
        v = RenderView()
        self._check(self.L.ws_read_render(self.ctx, mask, ctypes.byref(v)), "read_render")
        out = {}
        for k in names:
            ptr = getattr(v, k)
            out[k] = np.ctypeslib.as_array(ptr, (v.nP,)) if v.nP and ptr else np.zeros(0)
        return out

    # ---- balls (moving solids): (n, 9) doubles x y z vx vy vz r rho submerged
    def set_balls(self, balls):
        b = np.ascontiguousarray(np.asarray(balls, np.float64).reshape(-1, 9))
        self._check(self.L.ws_set_balls(self.ctx, len(b), b.ctypes.data_as(ctypes.c_void_p)), "set_balls")

    def balls(self):
        n = int(self.get("nBalls"))
        out = np.zeros((n, 9), np.float64)
        self._check(self.L.ws_get_balls(self.ctx, n, out.ctypes.data_as(ctypes.c_void_p)), "get_balls")
        return out

    # ---- the learned pressure warm start (include/watersim_learn.h)
    def net_params(self):
        n = self.L.ws_net_param_count()
        w = np.zeros(n, np.float32)
        self._check(self.L.ws_net_get_params(self.ctx, w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n),
                    "net_get_params")
        return w

    def set_net_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self._check(self.L.ws_net_set_params(self.ctx, w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(w)),
                    "net_set_params")

    def load_net_weights(self, path):
        """Trained weights as written by train_warmstart.py: the flat float32
        vector and, beside it, its .json record, whose "net_coarse" says on
        which grid the network was trained (1: the solver's grid coarsened
        2x per axis); the context is set to match."""
        self.set_net_params(np.fromfile(path, dtype=np.float32))
        meta = os.path.splitext(path)[0] + ".json"
        coarse = 0
        if os.path.exists(meta):
            import json
            with open(meta) as f:
                coarse = int(json.load(f).get("net_coarse", 0))
        self.set(netCoarse=coarse)

    def net_forward(self, inp):
        """inp [Z, Y, X, 5] -> out [Z, Y, X]"""
        inp = np.ascontiguousarray(inp, np.float32)
        Z, Y, X = inp.shape[:3]
        out = np.zeros((Z, Y, X), np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        self._check(self.L.ws_net_forward(self.ctx, Z, Y, X, inp.ctypes.data_as(fp), out.ctypes.data_as(fp)),
                    "net_forward")
        return out

    def net_grad(self, inp, target, mask):
        inp = np.ascontiguousarray(inp, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        mask = np.ascontiguousarray(mask, np.uint8)
        Z, Y, X = inp.shape[:3]
        g = np.zeros(self.L.ws_net_param_count(), np.float32)
        loss = ctypes.c_double()
        fp = ctypes.POINTER(ctypes.c_float)
        self._check(self.L.ws_net_grad(self.ctx, Z, Y, X, inp.ctypes.data_as(fp), target.ctypes.data_as(fp),
                                       mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), g.ctypes.data_as(fp),
                                       ctypes.byref(loss)), "net_grad")
        return loss.value, g

    def net_bank(self, capacity):
        self._check(self.L.ws_net_bank(self.ctx, capacity), "net_bank")

    def net_bank_count(self):
        n = ctypes.c_int()
        self._check(self.L.ws_net_bank_count(self.ctx, ctypes.byref(n)), "net_bank_count")
        return n.value

    def net_train(self, steps, batch, crop, lr, seed=1):
        loss = ctypes.c_double()
        self._check(self.L.ws_net_train(self.ctx, steps, batch, crop, lr, seed, ctypes.byref(loss)), "net_train")
        return loss.value

    def splat_density(self):
        self._check(self.L.ws_splat_density(self.ctx), "splat_density")

    def metaballs_build(self, tension, detail=float("nan")):
        """Metaballs.build(solver, tension, detail): (count, volume, pos, nrm,
        iso, lattice cells)."""
        cnt, vol = ctypes.c_long(), ctypes.c_double()
        iso, m = ctypes.c_double(), ctypes.c_int()
        self._check(self.L.ws_metaballs_build(self.ctx, tension, detail, None, None, 0, ctypes.byref(cnt),
                                              ctypes.byref(vol), ctypes.byref(iso), ctypes.byref(m)), "metaballs")
        n = cnt.value
        pos = np.zeros(max(n, 1) * 3, np.float32)
        nrm = np.zeros(max(n, 1) * 3, np.float32)
        self._check(self.L.ws_metaballs_build(self.ctx, tension, detail, pos.ctypes.data_as(ctypes.c_void_p),
                                              nrm.ctypes.data_as(ctypes.c_void_p), n, ctypes.byref(cnt),
                                              ctypes.byref(vol), ctypes.byref(iso), ctypes.byref(m)), "metaballs")
        return cnt.value, vol.value, pos[:n * 3], nrm[:n * 3], iso.value, m.value

    def mt_build(self, dens=None, dims=None, dx=None, iso=1.8, want_mesh=True):
        """MarchingTetrahedra.build on the context field (dens None) or a host
        field; the mesh lands in grow-only host buffers reused across calls
        (capacity doubling from 32768 vertices, as js/surface.js:53-64)."""
        nx, ny, nz = dims if dims else (self.nx, self.ny, self.nz)
        dx = self.dx if dx is None else dx
        cnt = ctypes.c_long()
        vol = ctypes.c_double()
        dptr = None
        if dens is not None:
            dens = np.ascontiguousarray(dens, np.float32)
            dptr = dens.ctypes.data_as(ctypes.c_void_p)
        if not want_mesh:
            self._check(self.L.ws_mt_build(self.ctx, dptr, nx, ny, nz, dx, iso, None, None, 0, ctypes.byref(cnt),
                                           ctypes.byref(vol)), "mt_build")
            return cnt.value, vol.value, None, None
        for _ in range(2):
            cap = 0 if self._mt_pos is None else len(self._mt_pos) // 3
            pp = self._mt_pos.ctypes.data_as(ctypes.c_void_p) if cap else None
            pn = self._mt_nrm.ctypes.data_as(ctypes.c_void_p) if cap else None
            self._check(self.L.ws_mt_build(self.ctx, dptr, nx, ny, nz, dx, iso, pp, pn, cap, ctypes.byref(cnt),
                                           ctypes.byref(vol)), "mt_build")
            n = cnt.value
            if n <= cap:
                return n, vol.value, self._mt_pos[:n * 3], self._mt_nrm[:n * 3]
            c = max(cap, 32768)
            while c < n:
                c *= 2
            self._mt_pos = np.zeros(c * 3, np.float32)
            self._mt_nrm = np.zeros(c * 3, np.float32)
        raise RuntimeError("mt_build: vertex count changed between calls")

    def mt_build_noalloc(self, dens_ptr, dims, dx, iso, pos_ptr=None, nrm_ptr=None, cap=0):
        """ws_mt_build_noalloc on device buffers given as integer addresses
        (dens_ptr None: the context's density); returns (count, volume)."""
        nx, ny, nz = dims
        cnt, vol = ctypes.c_long(), ctypes.c_double()
        self._check(self.L.ws_mt_build_noalloc(self.ctx, dens_ptr, nx, ny, nz, dx, iso, pos_ptr, nrm_ptr, cap,
                                               ctypes.byref(cnt), ctypes.byref(vol)), "mt_build_noalloc")
        return cnt.value, vol.value

    def mark_fluid(self, p0, p1):
        self._check(self.L.ws_mark_fluid(self.ctx, p0, p1), "mark_fluid")

    def sample_vel(self, pts):
        """sampleVel at the rows of pts (n x 3) with the device sampler."""
        xyz = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        out = np.empty_like(xyz)
        self._check(self.L.ws_sample_vel(self.ctx, xyz.ctypes.data_as(ctypes.c_void_p), len(xyz),
                                         out.ctypes.data_as(ctypes.c_void_p)), "sample_vel")
        return out

    def autotune_iso(self):
        out = ctypes.c_double()
        self._check(self.L.ws_autotune_iso(self.ctx, ctypes.byref(out)), "autotune_iso")
        return out.value

    def kinetic_energy(self):
        out = ctypes.c_double()
        self._check(self.L.ws_kinetic_energy(self.ctx, ctypes.byref(out)), "kinetic_energy")
        return out.value


class Batch:
    """Planets advanced together (ws_batch_*): each planet's stages on its
    own stream, all pressure solves in one launch, optionally replayed from
    a HIP graph. The solvers stay owned by the caller."""

    def __init__(self, solvers, graph=True, threads=1024):
        self.L = load()
        self.solvers = list(solvers)
        arr = (ctypes.c_void_p * len(self.solvers))(*[s.ctx.value for s in self.solvers])
        self.b = ctypes.c_void_p()
        rc = self.L.ws_batch_create(ctypes.byref(self.b), arr, len(self.solvers), int(bool(graph)), int(threads))
        if rc != 0:
            raise WatersimError("ws_batch_create failed (%d)" % rc)

    def run(self, substeps):
        self._check(self.L.ws_batch_run(self.b, int(substeps)), "batch run")

    def _check(self, rc, what):
        if rc != 0:
            msg = self.L.ws_batch_last_error(self.b)
            raise WatersimError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def set(self, **kw):
        for k, v in kw.items():
            self._check(self.L.ws_batch_set(self.b, k.encode(), float(v)), "batch set " + k)

    def get(self, k):
        v = ctypes.c_double()
        self._check(self.L.ws_batch_get(self.b, k.encode(), ctypes.byref(v)), "batch get " + k)
        return v.value

    def close(self):
        if self.b:
            self.L.ws_batch_destroy(self.b)
            self.b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_world(n, dx, coreR, depth, seed, target, bumpiness=1.0):
    """Planet + ocean fill on the host (ws_world_build); returns numpy arrays.
    Needs no GPU."""
    L = load()
    p = WorldParams(n, n, n, dx, coreR, depth, bumpiness, seed, target)
    w = WorldC()
    rc = L.ws_world_build(ctypes.byref(p), ctypes.byref(w))
    if rc != 0:
        raise WatersimError("ws_world_build failed (%d)" % rc)
    try:
        tn = w.tn
        uN, vN = (n + 1) * n * n, n * (n + 1) * n
        F = 3 * (n + 1) * n * n
        grav = np.ctypeslib.as_array(w.gravity, (F,)).copy()
        out = {
            "tn": tn, "dv": w.dv, "nP": w.nP, "capacity": w.capacity, "spacing": w.spacing,
            "particleVolume": w.particleVolume, "oceanR": w.oceanR, "coreR": w.coreR, "Rlo": w.Rlo,
            "Rhi": w.Rhi, "landFrac": w.landFrac, "nx": n, "dx": dx,
            "solid": np.ctypeslib.as_array(w.solid, (tn ** 3,)).copy(),
            "R": np.ctypeslib.as_array(w.R, (tn ** 3,)).copy(),
            "staticTypes": np.ctypeslib.as_array(w.staticTypes, (n ** 3,)).copy(),
            "gu": grav[:uN], "gv": grav[uN:uN + vN], "gw": grav[uN + vN:],
            "px": np.ctypeslib.as_array(w.px, (max(w.nP, 1),))[:w.nP].copy(),
            "py": np.ctypeslib.as_array(w.py, (max(w.nP, 1),))[:w.nP].copy(),
            "pz": np.ctypeslib.as_array(w.pz, (max(w.nP, 1),))[:w.nP].copy(),
        }
    finally:
        L.ws_world_free(ctypes.byref(w))
    return out


def planet_gravity(n, dx, oceanR):
    """ws_planet_gravity: (gu, gv, gw) of a cubic n^3 grid."""
    L = load()
    uN = (n + 1) * n * n
    g = np.empty(3 * uN, np.float32)
    if L.ws_planet_gravity(n, n, n, dx, oceanR, g.ctypes.data_as(ctypes.c_void_p)) != 0:
        raise WatersimError("ws_planet_gravity failed")
    return g[:uN], g[uN:2 * uN], g[2 * uN:]


def solver_for_world(world, capacity=None, Tamb=0.32):
    """A GPU context holding a built world, its fill at rest and gRamp = 0,
    as FluidSolver.resetWater leaves it (solver.js:718-730)."""
    n = world["nx"]
    s = Solver(n, n, n, world["dx"], capacity or world["capacity"])
    load_world(s, world, Tamb)
    return s


def load_world(s, world, Tamb=0.32):
    """Replaces the world and fill of an existing context of the same grid."""
    n, nP = world["nx"], world["nP"]
    if nP > s.capacity:
        raise WatersimError("world has %d particles, the context holds %d" % (nP, s.capacity))
    s.set(coreR=world["coreR"], oceanR=world["oceanR"], spacing=world["spacing"],
          particleVolume=world["particleVolume"], dv=world["dv"], tn=world["tn"])
    s.upload("staticTypes", world["staticTypes"])
    for k in ("gu", "gv", "gw"):
        s.upload(k, world[k])
    s.upload("terrainR", world["R"])
    s.upload("terrainSolid", world["solid"])
    s.set(Rlo=world["Rlo"], Rhi=world["Rhi"])
    s.set(nP=nP)
    s.upload("px", world["px"])
    s.upload("py", world["py"])
    s.upload("pz", world["pz"])
    z = np.zeros(nP, np.float32)
    for k in ("pvx", "pvy", "pvz", "pAir"):
        s.upload(k, z)
    s.upload("pT", np.full(nP, Tamb, np.float32))
    s.upload("pflag", np.zeros(nP, np.uint8))
    s.upload("pcool", np.zeros(nP, np.uint8))
    s.upload("q", np.zeros(n ** 3, np.float32))
    s.set(gRamp=0.0, simTime=0.0, heatAcc=0.0)
